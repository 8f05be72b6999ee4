"""Per-kernel timing of the wavefront render (rt_ctx_set_kernel_timing /
rt_ctx_kernel_stats): one closest-hit and one any-hit traversal launch per
sample and bounce, durations recorded, work counts split per kernel that
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
    assert ks["closest"]["launches"] == launches and ks["any"]["launches"] == launches
    assert ks["closest"]["ms"] > 0 and ks["any"]["ms"] > 0
    assert ks["closest"]["box_tests"] + ks["any"]["box_tests"] == c[3]
    assert ks["closest"]["tri_tests"] + ks["any"]["tri_tests"] == c[4]
    assert ks["closest"]["box_tests"] > 0 and ks["any"]["box_tests"] > 0
    # the stats reset on reading, and nothing is recorded once disabled
    rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts)
    ks2 = ctx.kernel_stats()
    assert all(v["launches"] == 0 and v["box_tests"] == 0 for v in ks2.values())
