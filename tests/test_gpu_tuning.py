"""Single-scene tuning: images, counts and hits do not depend on the tuning
(node layout, workgroup size, waves per SIMD, fetch chunk, stack ring);
rt_render_autotune stores its pick in the context's table under the
scene's size bucket; the table survives a save / load round trip, skips
other architectures' entries and rejects malformed files whole."""
import itertools

import numpy as np
import pytest

import rtamd
from oracle import oracle as O
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


@pytest.fixture(scope="module")
def scene(ctx):
    hm = scenes.soup_mesh(30_000, seed=21)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 63)
    mats, light, cam, opts = scenes.c2_setup(64, 48, 2)
    opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    return dict(hm=hm, dm=dm, bvh=b, mats=mats, md=rtamd.materials_tensor(mats), light=light, cam=cam, opts=opts)


def _render(ctx, sc):
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    rgb = rtamd.render(ctx, sc["bvh"], sc["dm"], sc["md"], len(sc["mats"]), sc["light"], sc["cam"], sc["opts"],
                       counters=cnt)
    torch.cuda.synchronize()
    return rgb.cpu().numpy(), cnt.cpu().numpy()


def test_render_and_trace_do_not_depend_on_tuning(ctx, scene):
    _say("single-scene renders and traces under every tuning axis")
    n, npix = scene["hm"].n, 64 * 48
    ref, ref_cnt = _render(ctx, scene)
    want, _ = O.render(scene["hm"], O.build(scene["hm"], 63), scene["mats"], scene["light"], scene["cam"],
                       scene["opts"])
    assert np.abs(ref - want).max() <= 1e-4
    orig, dirs = scenes.random_rays(5000, 22, spread=0.5)
    o_t, d_t = torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda()
    ref_hits = [scene["bvh"].trace(scene["dm"], o_t, d_t, m)[0].cpu().numpy() for m in (0, 1)]
    for layout, wg, waves, chunk, ring in itertools.product((2, 4, 5, 8), (64, 256), (1, 8), (16, 128), (16, 32)):
        t = rtamd.Tuning(ring, chunk, waves, layout, wg)
        ctx.set_tuning(n, npix, t)
        got, cnt = _render(ctx, scene)
        assert np.array_equal(got, ref), t.as_dict()
        assert np.array_equal(cnt[:3], ref_cnt[:3]), t.as_dict()
        ctx.set_tuning(n, 5000, t)
        for m in (0, 1):
            hits = scene["bvh"].trace(scene["dm"], o_t, d_t, m)[0]
            torch.cuda.synchronize()
            assert np.array_equal(hits.cpu().numpy(), ref_hits[m]), (m, t.as_dict())
    ctx.set_tuning(n, npix, None)
    ctx.set_tuning(n, 5000, None)
    assert not ctx.get_tuning(n, npix)[1]
    with pytest.raises(rtamd.RtError):
        ctx.set_tuning(n, npix, rtamd.Tuning(32, 16, 4, 5, 128))


def test_render_autotune_stores_its_pick(ctx, scene):
    _say("rt_render_autotune")
    n, npix = scene["hm"].n, 64 * 48
    ref, _ = _render(ctx, scene)
    best, ms, img = rtamd.render_autotune(ctx, scene["bvh"], scene["dm"], scene["md"], len(scene["mats"]),
                                          scene["light"], scene["cam"], scene["opts"])
    torch.cuda.synchronize()
    assert ms > 0 and best.node_layout in (2, 4, 5, 8) and best.wg_threads in (64, 256)
    assert np.array_equal(img.cpu().numpy(), ref)
    t, found = ctx.get_tuning(n, npix)
    assert found and t.as_dict() == best.as_dict()
    # the same bucket: 2^14 <= 30,000 < 2^15 triangles, 2^11 <= 3072 < 2^12 rays
    assert ctx.get_tuning(20_000, 2100)[1] and not ctx.get_tuning(40_000, npix)[1]
    got, _ = _render(ctx, scene)
    assert np.array_equal(got, ref)
    _say(f"autotune pick {best.as_dict()} {ms:.3f} ms")


def test_tuning_table_save_load(ctx, tmp_path):
    _say("tuning table save / load")
    c = rtamd.Context(0)
    t1, t2 = rtamd.Tuning(16, 64, 8, 8, 256), rtamd.Tuning(32, 32, 2, 4, 64)
    c.set_tuning(1_000_000, 1920 * 1080, t1)
    c.set_tuning(5_000, 256 * 256, t2)
    path = tmp_path / "tuning.txt"
    c.save_tuning(str(path))
    lines = [l for l in path.read_text().splitlines() if not l.startswith("#")]
    assert len(lines) == 2 and all(l.split()[0].startswith("gfx") for l in lines)
    d = rtamd.Context(0)
    d.load_tuning(str(path))
    assert d.get_tuning(1_000_000, 1920 * 1080)[0].as_dict() == t1.as_dict()
    assert d.get_tuning(5_000, 256 * 256)[0].as_dict() == t2.as_dict()
    # other architectures' entries are skipped
    other = tmp_path / "other.txt"
    other.write_text("# comment\ngfx000 3 3 16 16 4 2 64\n")
    e = rtamd.Context(0)
    e.load_tuning(str(other))
    assert not e.get_tuning(8, 8)[1]
    # a malformed line or an invalid tuning rejects the whole file
    arch = lines[0].split()[0]
    for bad in (f"{arch} 3 3 16 16 4 2 64\n{arch} 4 4 16 16\n", f"{arch} 3 3 16 16 4 7 64\n"):
        other.write_text(bad)
        with pytest.raises(rtamd.RtError) as err:
            e.load_tuning(str(other))
        assert err.value.status == rtamd.RT_ERR_INVALID
        assert not e.get_tuning(8, 8)[1]
    with pytest.raises(rtamd.RtError):
        e.load_tuning(str(tmp_path / "missing.txt"))
