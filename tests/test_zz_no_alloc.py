"""rt_render_noalloc (the render that allocates nothing, for callers that
cannot allocate device memory) against rt_render's wavefront path on the
same inputs, through the C-ABI: two calls in a row, the second smaller."""
import numpy as np
import pytest

import rtamd
from tests import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


def _case(ctx, hm, w, h, spp, depth, seed):
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, 30)
    mats, light, cam, _ = scenes.c1_setup(w, h, spp)
    mt = rtamd.materials_tensor(mats)
    cam.fstop = 6.0
    cam.focal_point = 4.0
    flags = rtamd.RT_RENDER_CLAMP | rtamd.RT_RENDER_WAVEFRONT
    want_c = torch.zeros(5, dtype=torch.int64, device="cuda")
    want = rtamd.render(ctx, bvh, dm, mt, len(mats), light, cam, rtamd.render_opts(spp, depth, seed, flags),
                        counters=want_c)
    # the no-alloc render writes into a poisoned buffer
    got = torch.full_like(want, float("nan"))
    got_c = torch.zeros(5, dtype=torch.int64, device="cuda")
    rtamd.render_noalloc(ctx, bvh, dm, mt, len(mats), light, cam,
                         rtamd.render_opts(spp, depth, seed, rtamd.RT_RENDER_CLAMP), got, counters=got_c)
    torch.cuda.synchronize()
    return want, got, want_c.cpu().numpy(), got_c.cpu().numpy(), bvh


def test_noalloc_render_matches_rt_render_twice_second_smaller(ctx):
    _say("noalloc call 1: icosphere freq 24 + quad (11,522 tris), 160x120, 3 spp, depth 1")
    want, got, wc, gc, bvh = _case(ctx, scenes.c1_mesh(24), 160, 120, 3, 1, 0)
    assert torch.equal(want, got)
    assert np.array_equal(wc[:3], gc[:3]) and wc[0] == 160 * 120 * 3 and wc[1] > 0
    bvh.close()
    _say("noalloc call 2: icosphere freq 8 + quad (1,282 tris), 48x40, 2 spp, depth 0")
    want, got, wc, gc, bvh = _case(ctx, scenes.c1_mesh(8), 48, 40, 2, 0, 7)
    assert torch.equal(want, got)
    assert np.array_equal(wc[:3], gc[:3]) and wc[0] == 48 * 40 * 2
    bvh.close()


def test_noalloc_render_rejects_sizes_beyond_its_scratch(ctx):
    _say("noalloc: oversize request is rejected")
    L = rtamd.lib()
    cap = int(L.rt_render_noalloc_max_bytes())
    assert cap == 48 << 20
    hm = scenes.c1_mesh(4)
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, 30)
    mats, light, cam, opts = scenes.c1_setup(1024, 1024, 1)
    mt = rtamd.materials_tensor(mats)
    rgb = torch.empty(1024 * 1024 * 3, dtype=torch.float32, device="cuda")
    with pytest.raises(rtamd.RtError) as e:
        rtamd.render_noalloc(ctx, bvh, dm, mt, len(mats), light, cam, opts, rgb)
    assert e.value.status == rtamd.RT_ERR_INVALID and "scratch" in str(e.value)
    bvh.close()
