"""Parity at the benchmark sizes: the 1,000,000-triangle soup of the
default workload (63-bit keys, 1920x1080, 16 spp thin lens) and the
9,996,980-triangle displaced sphere of the refit workload. The build and
refit must match the oracle bit for bit at full size; traversal ids exactly
on a ray sample; the full-frame render within 1e-4 on a pixel sample (the
oracle renders only those pixels)."""
import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from tests import scenes
from tests.test_gpu_parity import _compare_build, _trace_compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


@pytest.fixture(scope="module")
def soup1m():
    return scenes.soup_mesh(1_000_000)


def test_soup1m_build_bit_exact(ctx, soup1m):
    _say("full size: soup 1M build, 63-bit")
    dm = rtamd.DeviceMesh(soup1m)
    b = rtamd.BVH(ctx, dm, 63)
    _compare_build(b.export(), O.build(soup1m, 63, nthreads=16), soup1m, b)


def test_soup1m_trace_sample(ctx, soup1m):
    _say("full size: soup 1M closest/any trace, 20k rays")
    orig, dirs = scenes.random_rays(20_000, 31, spread=0.35)
    assert _trace_compare(ctx, soup1m, 63, orig, dirs) > 1000


def test_soup1m_full_frame_render_sample(ctx, soup1m):
    _say("full size: soup 1M render 1920x1080 x16 spp, oracle on every 211th pixel")
    dm = rtamd.DeviceMesh(soup1m)
    b = rtamd.BVH(ctx, dm, 63)
    mats, light, cam, opts = scenes.c2_setup(1920, 1080, 16)
    opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    rgb = rtamd.render(ctx, b, dm, rtamd.materials_tensor(mats), len(mats), light, cam, opts, counters=cnt)
    torch.cuda.synchronize()
    got = rgb.cpu().numpy().reshape(-1, 3)
    assert int(cnt[0]) == 1920 * 1080 * 16 and b.diagnostics() == (0, 0)
    # pixels in scan order; both images store row py at row H-1-py
    pixels = np.arange(0, 1920 * 1080, 211, dtype=np.int64)
    ref, counts = O.render(soup1m, O.build(soup1m, 63, nthreads=16), mats, light, cam, opts, pixels=pixels,
                           nthreads=16)
    assert counts[0] == pixels.size * 16
    px, py = pixels % 1920, pixels // 1920
    oi = (1080 - 1 - py) * 1920 + px
    d = np.abs(got[oi] - ref.reshape(-1, 3)[oi])
    assert d.max() <= 1e-4, d.max()


def test_sphere10m_build_and_refit_bit_exact(ctx):
    hm = rtamd.gen_icosphere(707, 1.0, (0.0, 0.0, -3.0), 0.05, 0x5EED0003)
    _say(f"full size: displaced sphere n={hm.n} build 30-bit + refit")
    assert hm.n == 9_996_980
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 30)
    ref = O.build(hm, 30, nthreads=16)
    _compare_build(b.export(), ref, hm, b)
    moved = rtamd.gen_jitter(hm, 1e-4, 0x5EED1000)
    dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
    b.refit(dm)
    got = b.export()
    tlo, thi = O.tri_aabb(moved)
    lo, hi = O.refit(ref["parent"], ref["left"], ref["right"], tlo, thi)
    assert np.array_equal(got["left"], ref["left"]) and np.array_equal(got["right"], ref["right"])
    assert np.array_equal(got["lo"].view(np.uint32), lo.view(np.uint32))
    assert np.array_equal(got["hi"].view(np.uint32), hi.view(np.uint32))
    assert b.diagnostics() == (0, 0)
