"""Per-kernel timing of the wavefront render (rt_ctx_set_kernel_timing /
rt_ctx_kernel_stats): one closest-hit traversal launch per sample and
bounce, one any-hit launch per sample and bounce but the last, where two
samples share one, durations recorded, work counts split per kernel that
add up to the render's own counters, and images unchanged by timing."""
import numpy as np
import pytest

import rtamd
from tests import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


@pytest.mark.parametrize("depth", [0, 2])
def test_kernel_stats_count_launches_and_work(ctx, depth):
    print(f"kernel timing depth={depth}", flush=True)
    hm = scenes.soup_mesh(20_000, seed=5)
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, 63)
    mats, light, cam, opts = scenes.c2_setup(96, 64, 3)
    opts.depth = depth
    opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    md = rtamd.materials_tensor(mats)
    ref = rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts).clone()
    ctx.kernel_stats()  # nothing recorded while disabled
    ctx.set_kernel_timing(True)
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    img = rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts, counters=cnt)
    ks = ctx.kernel_stats()
    ctx.set_kernel_timing(False)
    c = cnt.cpu().numpy()
    assert torch.equal(img, ref)
    launches = opts.spp * (depth + 1)
    # the last bounce's any-hit passes of samples 2j and 2j + 1 share a launch
    any_launches = opts.spp * depth + (opts.spp + 1) // 2
    assert ks["closest"]["launches"] == launches and ks["any"]["launches"] == any_launches
    assert ks["closest"]["ms"] > 0 and ks["any"]["ms"] > 0
    assert ks["closest"]["box_tests"] + ks["any"]["box_tests"] == c[3]
    assert ks["closest"]["tri_tests"] + ks["any"]["tri_tests"] == c[4]
    assert ks["closest"]["box_tests"] > 0 and ks["any"]["box_tests"] > 0
    # rays per kernel: every primary and bounce ray; the shadow rays whose
    # answer can change the image (the counters count every one cast)
    assert ks["closest"]["rays"] == c[0] + c[2]
    assert 0 < ks["any"]["rays"] <= c[1]
    # the stats reset on reading, and nothing is recorded once disabled
    rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts)
    ks2 = ctx.kernel_stats()
    assert all(v["launches"] == 0 and v["box_tests"] == 0 for v in ks2.values())


@pytest.mark.parametrize("depth", [0, 2])
def test_overlapped_schedule_matches_one_stream(ctx, depth):
    """The default schedule (a sample's last any-hit pass on a second stream,
    overlapping the next sample) gives the same image and ray counts as
    RT_RENDER_ONE_STREAM, over enough samples to reuse each parity buffer."""
    print(f"overlap vs one stream depth={depth}", flush=True)
    hm = scenes.soup_mesh(30_000, seed=9)
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, 63)
    mats, light, cam, opts = scenes.c2_setup(80, 48, 5)
    md = rtamd.materials_tensor(mats)
    out = []
    for extra in (0, rtamd.RT_RENDER_ONE_STREAM):
        o = rtamd.render_opts(5, depth, 3, opts.flags | rtamd.RT_RENDER_WAVEFRONT | extra)
        cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
        img = rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, o, counters=cnt)
        torch.cuda.synchronize()
        out.append((img.cpu().numpy(), cnt.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1][:3], out[1][1][:3])
