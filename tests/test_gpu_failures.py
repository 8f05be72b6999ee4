"""Failure paths of the C-ABI, forced from the host.

- Allocation failure: rt_ctx_fail_alloc makes the k-th device allocation
  fail. Every call that allocates (build, wavefront render workspace and
  shading table, batch, kernel timers) must return RT_ERR_OOM, hold no more
  device memory than before (rt_ctx_device_bytes), and leave the context
  and its other objects working. The reference exits the process instead
  (src/scene.cpp:71-81).
- hipGraph capture of a call that would have to allocate (a render larger
  than the context's workspace) or that gets a mesh of another size
  (rebuild, refit): RT_ERR_INVALID, nothing launched, the capture intact.
- NaN vertex coordinates: the build, refit, traces and render equal the
  oracle's (a NaN t, alpha or beta is not a hit, on both sides).
- Infinite vertex coordinates: the build equals the oracle's bit for bit,
  and traces terminate with valid answers (every reported hit is a hit of
  that ray by the oracle's triangle test); which hits are found is not
  specified, since the fp32 child tests are only conservative for finite
  boxes.
"""
import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from tests import scenes
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


def _status(fn):
    with pytest.raises(rtamd.RtError) as e:
        fn()
    return e.value.status


def test_allocation_failures_return_oom_without_leaks():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    _say("fault injection: each allocation of build, render, batch and timers fails in turn")
    c = rtamd.Context(0)
    assert c.device_bytes() == 0
    hm = scenes.soup_mesh(5000, seed=3)
    dm = rtamd.DeviceMesh(hm)
    # the BVH arena
    c.fail_alloc(1)
    assert _status(lambda: rtamd.BVH(c, dm, 30)) == rtamd.RT_ERR_OOM
    assert c.device_bytes() == 0
    b = rtamd.BVH(c, dm, 30)
    arena = c.device_bytes()
    assert arena > 0
    # the wavefront workspace (1st), then the shading table (2nd)
    mats, light, cam, opts = scenes.c2_setup(48, 32, 1)
    md = rtamd.materials_tensor(mats)
    wf = rtamd.render_opts(1, 0, 0, rtamd.RT_RENDER_WAVEFRONT)
    c.fail_alloc(1)
    assert _status(lambda: rtamd.render(c, b, dm, md, len(mats), light, cam, wf)) == rtamd.RT_ERR_OOM
    assert c.device_bytes() == arena
    c.fail_alloc(2)
    assert _status(lambda: rtamd.render(c, b, dm, md, len(mats), light, cam, wf)) == rtamd.RT_ERR_OOM
    held = c.device_bytes()
    assert held > arena  # the workspace stays with the context, to be reused
    img = rtamd.render(c, b, dm, md, len(mats), light, cam, wf).cpu().numpy()
    ref, _ = O.render(hm, O.build(hm, 30), mats, light, cam, opts)
    assert np.abs(img - ref).max() <= 1e-4
    held = c.device_bytes()
    # a batch makes three allocations (workspace, shading tables, build table)
    rgb = torch.zeros(48 * 32 * 3, dtype=torch.float32, device="cuda")
    item = dict(mesh=dm, bvh=b, mats=md, nmat=len(mats), light=light, cam=cam, opts=wf, rgb=rgb)
    for k in (1, 2, 3):
        c.fail_alloc(k)
        assert _status(lambda: rtamd.Batch(c, [item])) == rtamd.RT_ERR_OOM, k
        assert c.device_bytes() == held, k
    c.fail_alloc(0)
    batch = rtamd.Batch(c, [item])
    batch.render()
    torch.cuda.synchronize()
    assert torch.equal(rgb.cpu(), torch.from_numpy(img))
    batch.close()
    assert c.device_bytes() == held
    # kernel timers
    c.fail_alloc(1)
    assert _status(lambda: c.set_kernel_timing(True)) == rtamd.RT_ERR_OOM
    assert c.device_bytes() == held
    c.fail_alloc(0)
    # the BVH that was built before the failures still traces like the oracle's
    orig, dirs = scenes.random_rays(2000, 9)
    tri = b.trace(dm, torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda(), rtamd.RT_TRACE_CLOSEST)[0]
    torch.cuda.synchronize()
    assert np.array_equal(tri.cpu().numpy(), O.trace(hm, O.build(hm, 30), orig, dirs)[0])
    b.close()
    assert c.device_bytes() == held - arena
    c.close()


def test_capture_that_needs_allocation_or_other_sizes_is_invalid(ctx):
    _say("graph capture: a larger render and a mesh of another size return RT_ERR_INVALID")
    hm = scenes.soup_mesh(4000, seed=5)
    dm = rtamd.DeviceMesh(hm)
    other = rtamd.DeviceMesh(scenes.soup_mesh(3000, seed=6))
    b = rtamd.BVH(ctx, dm, 30)
    mats, light, _, _ = scenes.c2_setup(32, 32, 1)
    md = rtamd.materials_tensor(mats)
    wf = rtamd.render_opts(1, 0, 0, rtamd.RT_RENDER_WAVEFRONT)
    small = rtamd.camera(32, 32, fov_deg=40.0, fstop=20.0, focal_point=1.5)
    # the workspace a 32x32 render needs, or a larger one from earlier tests
    rtamd.render(ctx, b, dm, md, len(mats), light, small, wf)
    torch.cuda.synchronize()
    big = rtamd.camera(256, 256, fov_deg=40.0, fstop=20.0, focal_point=1.5)
    rgb = torch.empty(256 * 256 * 3, dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    for what, fn in (("render", lambda: rtamd.render(ctx, b, dm, md, len(mats), light, big, wf, rgb=rgb)),
                     ("rebuild", lambda: b.rebuild(other)), ("refit", lambda: b.refit(other))):
        g = torch.cuda.CUDAGraph()
        with pytest.raises(rtamd.RtError) as e:
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    fn()
        assert e.value.status == rtamd.RT_ERR_INVALID, (what, e.value)
        torch.cuda.synchronize()
    # still usable: the same render eagerly, then captured and replayed
    rtamd.render(ctx, b, dm, md, len(mats), light, big, wf, rgb=rgb)
    torch.cuda.synchronize()
    want = rgb.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            rtamd.render(ctx, b, dm, md, len(mats), light, big, wf, rgb=rgb)
    torch.cuda.current_stream().wait_stream(side)
    for _ in range(2):
        rgb.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(rgb, want)
    del rgb, want


def _nan_mesh(n, seed):
    """A soup where 40 triangles have one vertex with a NaN x and 25 have a
    NaN y on all three vertices (a NaN box on y)."""
    hm = scenes.soup_mesh(n, seed=seed)
    rng = np.random.default_rng(seed)
    tris = rng.choice(n, 65, replace=False)
    vx, vy = hm.vx.copy(), hm.vy.copy()
    for t in tris[:40]:
        vx[hm.idx[3 * t + 1]] = np.nan
    for t in tris[40:]:
        vy[hm.idx[3 * t: 3 * t + 3]] = np.nan
    return rtamd.HostMesh(vx, vy, hm.vz, hm.nx, hm.ny, hm.nz, hm.idx, hm.mat_id)


def _boxes_equal(a, b):
    a, b = a.reshape(-1), b.reshape(-1)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def _export_matches(got, ref):
    for k in ("codes", "ids", "parent", "left", "right"):
        assert np.array_equal(got[k].reshape(-1), ref[k].astype(got[k].dtype).reshape(-1)), k
    for k in ("lo", "hi"):
        assert _boxes_equal(got[k], ref[k]), k


@pytest.mark.parametrize("bits", [30, 63])
def test_nan_vertices_match_oracle(ctx, bits):
    _say(f"NaN vertices: build, refit, traces, render vs the oracle ({bits}-bit keys)")
    hm = _nan_mesh(20_000, 11)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, bits)
    ref = O.build(hm, bits)
    _export_matches(b.export(), ref)
    moved = rtamd.gen_jitter(hm, 1e-3, 77)
    dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
    b.refit(dm)
    tlo, thi = O.tri_aabb(moved)
    lo, hi = O.refit(ref["parent"], ref["left"], ref["right"], tlo, thi)
    ref_moved = dict(ref, lo=lo, hi=hi)
    _export_matches(b.export(), ref_moved)
    orig, dirs = scenes.random_rays(6000, 12, spread=0.5)
    # rays aimed at the NaN triangles' finite vertices too
    for any_hit in (False, True):
        tri = b.trace(dm, torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda(),
                      rtamd.RT_TRACE_ANY if any_hit else rtamd.RT_TRACE_CLOSEST)[0]
        torch.cuda.synchronize()
        rtri = O.trace(moved, ref_moved, orig, dirs, any_hit=any_hit)[0]
        assert np.array_equal(tri.cpu().numpy(), rtri), any_hit
        assert (rtri >= 0).sum() > 50
    assert b.diagnostics() == (0, 0)
    mats, light, cam, opts = scenes.c2_setup(64, 48, 2)
    _render_compare(ctx, moved, bits, mats, light, cam, opts)


def test_inf_vertices_build_exact_and_traces_terminate(ctx):
    _say("infinite vertices: exact build, terminating traces with valid hits")
    hm = scenes.soup_mesh(20_000, seed=13)
    vx, vz = hm.vx.copy(), hm.vz.copy()
    rng = np.random.default_rng(13)
    for k, t in enumerate(rng.choice(hm.n, 12, replace=False)):
        (vx if k % 2 else vz)[hm.idx[3 * t + k % 3]] = np.inf if k % 4 < 2 else -np.inf
    hm = rtamd.HostMesh(vx, hm.vy, vz, hm.nx, hm.ny, hm.nz, hm.idx, hm.mat_id)
    dm = rtamd.DeviceMesh(hm)
    for bits in (30, 63):
        b = rtamd.BVH(ctx, dm, bits)
        _export_matches(b.export(), O.build(hm, bits))
        orig, dirs = scenes.random_rays(4000, 14, spread=0.5)
        for mode in (rtamd.RT_TRACE_CLOSEST, rtamd.RT_TRACE_ANY):
            tri = b.trace(dm, torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda(), mode)[0]
            torch.cuda.synchronize()
            tri = tri.cpu().numpy()
            assert ((tri >= -1) & (tri < hm.n)).all()
            hit = np.nonzero(tri >= 0)[0]
            assert hit.size > 50
            for r in hit[:200]:
                assert O.intersect_triangle(hm, int(tri[r]), orig[r], dirs[r])[0], (mode, r)
        assert b.diagnostics() == (0, 0)
        b.close()
