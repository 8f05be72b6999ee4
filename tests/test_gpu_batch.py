"""Batched rendering (rt_batch_*): every scene's image equals the single-scene
render bit for bit, for every tuning, eagerly and replayed from a hipGraph."""
import itertools

import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from tests import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


def _make(ctx, specs, bits=63):
    out = []
    for sp in specs:
        dm = rtamd.DeviceMesh(sp["hm"])
        b = rtamd.BVH(ctx, dm, bits)
        md = rtamd.materials_tensor(sp["mats"])
        rgb = torch.zeros(sp["cam"].width * sp["cam"].height * 3, dtype=torch.float32, device="cuda")
        out.append(dict(mesh=dm, bvh=b, mats=md, nmat=len(sp["mats"]), light=sp["light"], cam=sp["cam"],
                        opts=sp["opts"], rgb=rgb, hm=sp["hm"]))
    return out


def _single(ctx, sc):
    o = rtamd.render_opts(sc["opts"].spp, sc["opts"].depth, sc["opts"].seed,
                          sc["opts"].flags | rtamd.RT_RENDER_WAVEFRONT)
    return rtamd.render(ctx, sc["bvh"], sc["mesh"], sc["mats"], sc["nmat"], sc["light"], sc["cam"], o)


def test_batch_matches_single_and_oracle(ctx):
    print("batch vs single", flush=True)
    specs = scenes.batch_scenes(6, 11, tri_range=(500, 30_000), resolutions=(32, 48, 64), spp_range=(1, 3))
    specs[2]["opts"].depth = 2
    specs[4]["cam"].width = 40  # non-square, not a multiple of 16
    specs[4]["cam"].height = 24
    sc = _make(ctx, specs)
    batch = rtamd.Batch(ctx, sc)
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    batch.render(counters=cnt)
    torch.cuda.synchronize()
    tot = np.zeros(3, np.int64)
    for k, s in enumerate(sc):
        single = _single(ctx, s)
        torch.cuda.synchronize()
        got = s["rgb"].cpu().numpy()
        assert np.array_equal(got, single.cpu().numpy()), f"scene {k} differs from single render"
        ref, counts = O.render(s["hm"], O.build(s["hm"], 63), specs[k]["mats"], s["light"], s["cam"], s["opts"])
        assert np.abs(got - ref).max() <= 1e-4, f"scene {k} differs from the oracle"
        tot += counts
    c = cnt.cpu().numpy()
    assert list(c[:3]) == list(tot)
    assert batch.diagnostics() == 0


_EXPORT_KEYS = ("codes", "ids", "parent", "left", "right", "lo", "hi", "tri_lo", "tri_hi")


@pytest.mark.parametrize("bits", [30, 63])
def test_batch_rebuild_matches_single_builds(ctx, bits):
    """rt_batch_rebuild builds every scene in shared launches (one per build
    stage): each BVH equals a single-scene build of the moved positions, bit
    for bit, and its image equals a single-scene render. Sizes cover one
    triangle, a few, chunk edges (512 leaves per climb chunk) and scenes
    that need several top-level passes."""
    sizes = [1, 2, 3, 511, 513, 1025, 4097, 70_000, 250_000]
    print(f"segmented batch rebuild bits={bits} sizes={sizes}", flush=True)
    specs = []
    for k, n in enumerate(sizes):
        sp = scenes.batch_scenes(1, 20 + k, tri_range=(n, n + 1), resolutions=(32,), spp_range=(1, 1))[0]
        specs.append(sp)
    sc = _make(ctx, specs, bits)
    batch = rtamd.Batch(ctx, sc)
    moved = [rtamd.gen_soup(s["mesh"].n, 0xBEEF + k) for k, s in enumerate(sc)]
    for s, hm in zip(sc, moved):
        s["mesh"].set_positions(*(torch.from_numpy(a).cuda() for a in (hm.vx, hm.vy, hm.vz)))
    batch.rebuild()
    batch.render()
    torch.cuda.synchronize()
    for k, (s, hm) in enumerate(zip(sc, moved)):
        got = s["bvh"].export()
        ref_bvh = rtamd.BVH(ctx, rtamd.DeviceMesh(hm), bits)
        ref = ref_bvh.export()
        for key in _EXPORT_KEYS:
            a, r = got[key], ref[key]
            if a.dtype == np.float32:
                a, r = a.view(np.uint32), r.view(np.uint32)
            assert np.array_equal(a, r), f"scene {k} (n={sizes[k]}): {key} differs"
        assert s["bvh"].diagnostics() == (0, 0)
        single = _single(ctx, s)
        torch.cuda.synchronize()
        assert np.array_equal(s["rgb"].cpu().numpy(), single.cpu().numpy()), f"scene {k} image differs"
    b0 = sc[6]["bvh"].export()
    ref = O.build(moved[6], bits)
    assert np.array_equal(b0["parent"][1:], ref["parent"][1:]), "parent links differ from the oracle"
    for key in ("codes", "ids", "left", "right"):
        assert np.array_equal(b0[key], ref[key]), f"{key} differs from the oracle"
    assert batch.diagnostics() == 0


def test_batch_items_sharing_a_bvh(ctx):
    """Two cameras on one scene: items 0 and 2 share a BVH and mesh. A batch
    rebuild builds that BVH once (two concurrent builds of one BVH would
    race), and every item's image equals its single-scene render."""
    print("batch items sharing a BVH", flush=True)
    specs = scenes.batch_scenes(2, 30, tri_range=(20_000, 40_000), resolutions=(48,), spp_range=(1, 2))
    sc = _make(ctx, specs)
    third = dict(sc[0])
    third["cam"] = scenes.c2_setup(64, 40, 1)[2]
    third["opts"] = rtamd.render_opts(2, 1, 7, sc[0]["opts"].flags)
    third["rgb"] = torch.zeros(64 * 40 * 3, dtype=torch.float32, device="cuda")
    sc.append(third)
    batch = rtamd.Batch(ctx, sc)
    for path in ("segmented", "per-scene"):
        if path == "per-scene":
            # a batch whose BVHs keep layout 8 nodes rebuilds them in per-scene launch chains
            batch.set_tuning(rtamd.Tuning(32, 64, 4, 8))
            batch.render()
        moved = rtamd.gen_soup(sc[0]["mesh"].n, 0xF00D if path == "segmented" else 0xF00E)
        sc[0]["mesh"].set_positions(*(torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)))
        batch.rebuild()
        batch.render()
        torch.cuda.synchronize()
        got = sc[0]["bvh"].export()
        ref = rtamd.BVH(ctx, rtamd.DeviceMesh(moved), 63).export()
        for key in _EXPORT_KEYS:
            a, r = got[key], ref[key]
            if a.dtype == np.float32:
                a, r = a.view(np.uint32), r.view(np.uint32)
            assert np.array_equal(a, r), f"{path}: shared BVH {key} differs"
        assert sc[0]["bvh"].diagnostics() == (0, 0)
        for k, s in enumerate(sc):
            single = _single(ctx, s)
            torch.cuda.synchronize()
            assert np.array_equal(s["rgb"].cpu().numpy(), single.cpu().numpy()), f"{path}: item {k} differs"
    assert batch.diagnostics() == 0


def test_batch_of_many_scenes_matches_single(ctx):
    """80 scenes: the shadow-ray sort key keeps 7 bits for the scene and a
    16 x 16 light-cell grid (64 scenes or fewer keep the 32 x 32 grid)."""
    print("batch of 80 scenes vs single", flush=True)
    specs = scenes.batch_scenes(80, 14, tri_range=(300, 4_000), resolutions=(16, 24, 32), spp_range=(1, 2))
    specs[7]["opts"].depth = 1
    sc = _make(ctx, specs)
    batch = rtamd.Batch(ctx, sc)
    batch.render()
    torch.cuda.synchronize()
    for k, s in enumerate(sc):
        single = _single(ctx, s)
        torch.cuda.synchronize()
        assert np.array_equal(s["rgb"].cpu().numpy(), single.cpu().numpy()), f"scene {k} differs from single render"
    assert batch.diagnostics() == 0


def test_batch_outputs_do_not_depend_on_tuning(ctx):
    print("batch tuning invariance", flush=True)
    specs = scenes.batch_scenes(5, 12, tri_range=(1000, 60_000), resolutions=(64, 96), spp_range=(1, 2))
    sc = _make(ctx, specs)
    batch = rtamd.Batch(ctx, sc)
    ref = None
    for layout, ring, chunk, waves, wg in itertools.product((2, 4, 5, 8), (16, 32), (16, 256), (1, 4, 8), (64, 256)):
        t = rtamd.Tuning(ring, chunk, waves, layout, wg)
        batch.set_tuning(t)
        for s in sc:
            s["rgb"].zero_()
        batch.render()
        torch.cuda.synchronize()
        imgs = [s["rgb"].cpu().numpy() for s in sc]
        if ref is None:
            ref = imgs
        else:
            for a, b in zip(ref, imgs):
                assert np.array_equal(a, b), t.as_dict()
    best, ms = batch.autotune()
    assert ms > 0 and best.stack_ring in (16, 32)
    batch.render()
    torch.cuda.synchronize()
    for a, s in zip(ref, sc):
        assert np.array_equal(a, s["rgb"].cpu().numpy())


def test_batch_graph_replay_matches_eager(ctx):
    print("batch graph replay", flush=True)
    specs = scenes.batch_scenes(4, 13, tri_range=(2000, 50_000), resolutions=(64, 128), spp_range=(1, 3))
    sc = _make(ctx, specs)
    batch = rtamd.Batch(ctx, sc)
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    batch.rebuild()
    batch.render(counters=cnt)  # warm-up outside capture
    torch.cuda.synchronize()
    eager = [s["rgb"].cpu().numpy().copy() for s in sc]
    eager_cnt = cnt.cpu().numpy().copy()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            cnt.zero_()
            batch.rebuild()
            batch.render(counters=cnt)
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        for s in sc:
            s["rgb"].fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, s in enumerate(sc):
            assert np.array_equal(eager[k], s["rgb"].cpu().numpy()), f"replay {rep} scene {k}"
        c = cnt.cpu().numpy()
        # ray counts are exact; the box and triangle test counts depend on
        # how a wave batches its leaf tests (closest-hit pruning), so they
        # agree only closely
        assert np.array_equal(c[:3], eager_cnt[:3])
        assert np.allclose(c[3:], eager_cnt[3:], rtol=0.02)
