"""Every configuration at its full size, against the oracle.

C1: the 20,482-triangle icosphere scene, 512x512 at 1 spp, every pixel.
C3: the 9,996,980-triangle displaced sphere after one jittered refit,
    1024x1024 at 4 spp with diffuse bounces to depth 4, on a pixel sample.
C4: the 64-scene batch (soups of 1K..2M triangles, 256^2..1024^2, 1..8 spp):
    batched images (eager and replayed from a hipGraph) equal the
    per-scene launch loop bit for bit, and three scenes match the oracle on
    a pixel sample.
C5: the 200,000-triangle differentiable scene, 512x512 at 16 spp, depth 2:
    the forward and its path records against the oracle on a pixel sample,
    and the albedo and light gradients against central finite differences
    of the GPU forward (1e-6 relative; the image is polynomial in both).
The oracle renders only the sampled pixels (its traversal is the
reference's: every box the line enters).
"""
import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from rtamd import diff, workloads
from tests.test_gpu_parity import _render_compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


def _sample(cam, stride, offset=0):
    """Launch pixels py*W+px of every stride-th pixel, and their rows in the
    row-flipped image (kernels/raytracer.cu:106)."""
    W, H = cam.width, cam.height
    pix = np.arange(offset % stride, W * H, stride, dtype=np.int64)
    px, py = pix % W, pix // W
    return pix, (H - 1 - py) * W + px


def test_c1_full_frame_512_1spp(ctx):
    _say("C1: icosphere freq 32 + quad, 512x512 x1 spp, megakernel and wavefront vs the oracle")
    hm = workloads.c1_mesh(32)
    assert hm.n == 20_482
    mats, light, cam, opts = workloads.c1_setup(512, 512, 1)
    got, d = _render_compare(ctx, hm, 30, mats, light, cam, opts)
    assert got.max() > 0.1 and (got == np.float32(0.9)).any()  # lit sphere, ground and background


def test_c3_refit_then_depth4_render(ctx):
    hm = workloads.sphere10m_mesh()
    _say(f"C3: displaced sphere n={hm.n}: build, jittered refit, 1024x1024 x4 spp depth 4")
    assert hm.n == 9_996_980
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 30)
    moved = workloads.sphere10m_frame(hm, 0)
    dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
    b.refit(dm)
    mats, light, cam, opts = workloads.sphere10m_setup()
    opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    rgb = rtamd.render(ctx, b, dm, rtamd.materials_tensor(mats), len(mats), light, cam, opts, counters=cnt)
    torch.cuda.synchronize()
    got = rgb.cpu().numpy().reshape(-1, 3)
    c = cnt.cpu().numpy()
    assert c[0] == 1024 * 1024 * 4 and c[2] > 100_000 and b.diagnostics() == (0, 0)
    # the oracle: the 10M build on the original positions, refit on the moved ones
    ref = O.build(hm, 30, nthreads=16)
    tlo, thi = O.tri_aabb(moved)
    lo, hi = O.refit(ref["parent"], ref["left"], ref["right"], tlo, thi)
    ref_moved = dict(ref, lo=lo, hi=hi)
    pix, oi = _sample(cam, 1009)
    want, counts = O.render(moved, ref_moved, mats, light, cam, opts, pixels=pix, nthreads=16)
    assert counts[0] == pix.size * 4 and counts[2] > 0
    d = np.abs(got[oi] - want.reshape(-1, 3)[oi])
    _say(f"C3: {pix.size} sampled pixels, {int(counts[2])} bounce rays, max diff {d.max():.3g}")
    assert d.max() <= 1e-4, d.max()
    b.close()


def test_c4_full_batch_matches_loop_graph_and_oracle(ctx):
    specs = workloads.batch_scenes(64)
    _say(f"C4: 64 scenes, {sum(s['hm'].n for s in specs)} triangles, batched vs per-scene loop")
    sc = []
    for sp in specs:
        dm = rtamd.DeviceMesh(sp["hm"])
        sp["opts"].flags |= rtamd.RT_RENDER_WAVEFRONT
        sc.append(dict(mesh=dm, bvh=rtamd.BVH(ctx, dm, 63), mats=rtamd.materials_tensor(sp["mats"]),
                       nmat=len(sp["mats"]), light=sp["light"], cam=sp["cam"], opts=sp["opts"],
                       rgb=torch.zeros(sp["cam"].width * sp["cam"].height * 3, dtype=torch.float32, device="cuda")))
    loop = []
    for s in sc:
        loop.append(rtamd.render(ctx, s["bvh"], s["mesh"], s["mats"], s["nmat"], s["light"], s["cam"], s["opts"]))
    batch = rtamd.Batch(ctx, sc)
    batch.rebuild()
    batch.render()
    torch.cuda.synchronize()
    for k, (s, want) in enumerate(zip(sc, loop)):
        assert torch.equal(s["rgb"], want), f"scene {k}: batched image differs from the per-scene render"
    # replayed from a hipGraph, with eager work between replays
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            batch.rebuild()
            batch.render()
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(2):
        for s in sc:
            s["rgb"].zero_()
        g.replay()
        torch.cuda.synchronize()
        for k, (s, want) in enumerate(zip(sc, loop)):
            assert torch.equal(s["rgb"], want), f"replay {rep}, scene {k} differs"
    assert batch.diagnostics() == 0
    # three scenes against the oracle: the smallest, the median and the largest
    order = np.argsort([s["mesh"].n for s in sc])
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        s, sp = sc[k], specs[k]
        pix, oi = _sample(s["cam"], max(1, s["cam"].width * s["cam"].height // 500) | 1, offset=k)
        want, counts = O.render(sp["hm"], O.build(sp["hm"], 63, nthreads=16), sp["mats"], s["light"], s["cam"],
                                s["opts"], pixels=pix, nthreads=16)
        got = s["rgb"].cpu().numpy().reshape(-1, 3)
        d = np.abs(got[oi] - want.reshape(-1, 3)[oi])
        _say(f"C4 scene {k}: n={s['mesh'].n} {s['cam'].width}^2 x{s['opts'].spp} spp, {pix.size} px, "
             f"max diff {d.max():.3g}")
        assert d.max() <= 1e-4, (k, d.max())
    batch.close()
    for s in sc:
        s["bvh"].close()


def test_c5_diff200k_forward_oracle_and_gradient_fd(ctx):
    hm = workloads.diff200k_mesh()
    _say(f"C5: differentiable render n={hm.n}, 512x512 x16 spp depth 2")
    assert hm.n == 200_000
    dm = rtamd.DeviceMesh(hm)
    light, cam = workloads.diff200k_setup()
    spp, depth, seed = 16, 2, 1
    R = diff.DiffRenderer(ctx, dm, light, cam, spp, depth, seed=seed)
    _, pert = workloads.diff200k_albedo(hm.n)
    albedo = torch.tensor(pert, dtype=torch.float64, device="cuda")
    s = 10.0
    img = R.forward(albedo, s).clone()
    rec = R.records.cpu().numpy().reshape(cam.width * cam.height, spp, depth + 1, 2)
    torch.cuda.synchronize()
    # the oracle on a pixel sample: one material per triangle, light colour x s
    m = rtamd.HostMesh(hm.vx, hm.vy, hm.vz, hm.nx, hm.ny, hm.nz, hm.idx, np.arange(hm.n, dtype=np.int32))
    mats = [rtamd.material(kd=tuple(a)) for a in pert]
    L = rtamd.light(position=tuple(light.position), color=tuple(s * c for c in light.color), radius=light.radius,
                    c0=light.c0, c1=light.c1, c2=light.c2)
    opts = rtamd.render_opts(spp, depth=depth, seed=seed, flags=0)
    pix, oi = _sample(cam, 499)
    out = O.render(m, O.build(m, 30, nthreads=16), mats, L, cam, opts, pixels=pix, nthreads=16, fp64=True,
                   records=True)
    want, rec_ref = out[2].reshape(-1, 3), out[3]
    got = img.cpu().numpy().reshape(-1, 3)
    d = np.abs(got[oi] - want[oi])
    _say(f"C5: {pix.size} sampled pixels, max forward diff {d.max():.3g}")
    assert d.max() <= 1e-4, d.max()
    ids = R.bvh.export()["ids"].astype(np.int64)
    r = rec[pix].copy()
    for c in (0, 1):
        v = r[..., c]
        v[v >= 0] = ids[v[v >= 0]]
    assert np.array_equal(r, rec_ref[pix]), "path records differ from the oracle's"
    # gradients: L = sum(w * image)
    rng = np.random.default_rng(5)
    w = torch.tensor(rng.normal(size=R.shape), dtype=torch.float64, device="cuda")
    g_alb, g_light, _ = R.backward(albedo, s, w)
    torch.cuda.synchronize()
    loss = lambda a, ss: float((R.forward(a, ss) * w).sum())
    h = 1e-3
    fd = (loss(albedo, s + h) - loss(albedo, s - h)) / (2 * h)
    assert abs(float(g_light) - fd) <= 1e-6 * abs(fd) + 1e-9, (float(g_light), fd)
    v = torch.tensor(rng.normal(size=(hm.n, 3)), dtype=torch.float64, device="cuda")
    h = 1e-4
    fd = (loss(albedo + h * v, s) - loss(albedo - h * v, s)) / (2 * h)
    pred = float((g_alb * v).sum())
    assert abs(pred - fd) <= 1e-6 * abs(fd) + 1e-9, (pred, fd)
    ga = g_alb.cpu().numpy()
    for t, c in zip(*np.unravel_index(np.argsort(-np.abs(ga).ravel())[:4], ga.shape)):
        e = torch.zeros_like(albedo)
        e[t, c] = 1.0
        fd = (loss(albedo + 1e-4 * e, s) - loss(albedo - 1e-4 * e, s)) / 2e-4
        assert abs(ga[t, c] - fd) <= 1e-6 * abs(fd) + 1e-9, (t, c, ga[t, c], fd)
    _say(f"C5: light and albedo gradients agree with central FD (dL/ds {float(g_light):.6g})")
