"""GPU parity: librtamd (HIP, gfx950) against the FP64 CPU oracle.

Build outputs (Morton codes, sorted ids, topology, parent links, node and
triangle AABBs, scene bounds) must match bit for bit; hit triangle ids
must match exactly with t / barycentrics within 1e-5 relative; rendered
images within 1e-4 per channel.
"""
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


def _compare_build(got, ref, hm, bvh=None):
    if bvh is not None:
        assert bvh.diagnostics() == (0, 0)
    assert np.array_equal(got["codes"], ref["codes"]), "morton codes / sort order differ"
    assert np.array_equal(got["ids"], ref["ids"]), "sorted ids differ"
    assert np.array_equal(got["parent"][1:], ref["parent"][1:]), "parent links differ"
    assert got["parent"][0] == -1
    assert np.array_equal(got["left"], ref["left"]), "left children differ"
    assert np.array_equal(got["right"], ref["right"]), "right children differ"
    assert np.array_equal(got["lo"].view(np.uint32), ref["lo"].view(np.uint32)), "node AABB min differ"
    assert np.array_equal(got["hi"].view(np.uint32), ref["hi"].view(np.uint32)), "node AABB max differ"
    tlo, thi = O.tri_aabb(hm)
    assert np.array_equal(got["tri_lo"].view(np.uint32), tlo.view(np.uint32))
    assert np.array_equal(got["tri_hi"].view(np.uint32), thi.view(np.uint32))


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("scene", ["c1", "soup", "displaced"])
def test_build_parity(ctx, bits, scene):
    hm = {"c1": lambda: scenes.c1_mesh(32), "soup": lambda: scenes.soup_mesh(150_000),
          "displaced": lambda: scenes.displaced_mesh(60)}[scene]()
    _say(f"build parity scene={scene} bits={bits} n={hm.n}")
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, bits)
    got = b.export()
    ref = O.build(hm, bits)
    _compare_build(got, ref, hm, b)
    mn, mx = b.scene_bounds()
    assert np.array_equal(mn, ref["mn"]) and np.array_equal(mx, ref["mx"])


@pytest.mark.parametrize("n", [1, 2, 3, 17, 511, 512, 513, 1023, 1024, 1025, 4095, 4096, 4097, 70_000])
def test_build_sizes(ctx, n):
    _say(f"build sizes n={n}")
    hm = scenes.soup_mesh(n, seed=1234 + n)
    dm = rtamd.DeviceMesh(hm)
    for bits in (30, 63):
        b = rtamd.BVH(ctx, dm, bits)
        got = b.export()
        ref = O.build(hm, bits)
        _compare_build(got, ref, hm, b)


@pytest.mark.parametrize("bits", [30, 63])
def test_build_karras_search_path(ctx, bits):
    """The per-node span/split search + atomic refit path gives the same BVH."""
    _say(f"karras search path bits={bits}")
    hm = scenes.soup_mesh(50_000, seed=99)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, bits, flags=rtamd.RT_BUILD_KARRAS_SEARCH)
    _compare_build(b.export(), O.build(hm, bits), hm)


@pytest.mark.parametrize("flags", [rtamd.RT_BUILD_DEFAULT, rtamd.RT_BUILD_KARRAS_SEARCH])
def test_refit_parity(ctx, flags):
    """Refit after vertex jitter: same topology, boxes = oracle refit of the moved mesh."""
    _say(f"refit parity flags={flags}")
    hm = scenes.displaced_mesh(40)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 63, flags=flags)
    ref = O.build(hm, 63)
    for frame in range(3):
        moved = rtamd.gen_jitter(hm, 1e-3, 1000 + frame)
        dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
        b.refit(dm)
        got = b.export()
        tlo, thi = O.tri_aabb(moved)
        lo, hi = O.refit(ref["parent"], ref["left"], ref["right"], tlo, thi)
        assert np.array_equal(got["left"], ref["left"]) and np.array_equal(got["right"], ref["right"])
        assert np.array_equal(got["parent"][1:], ref["parent"][1:])
        assert np.array_equal(got["lo"].view(np.uint32), lo.view(np.uint32))
        assert np.array_equal(got["hi"].view(np.uint32), hi.view(np.uint32))
        assert b.diagnostics() == (0, 0)


@pytest.mark.parametrize("n", [300, 30_000])
def test_refit_far_motion_traces(ctx, n):
    """Refit after every triangle moved far (another soup with the same
    triangle count): the tree keeps the old topology, so every box and every
    quantized 4-wide node changes. Closest and any hits on the refit BVH
    must equal the oracle's on the same topology with refit boxes; a wide
    node the refit left stale would hide hits. n = 300 fits one climb chunk;
    30,000 needs the top-level passes."""
    _say(f"refit far motion n={n}")
    hm = scenes.soup_mesh(n, seed=41)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 63)
    ref = O.build(hm, 63)
    moved = rtamd.gen_soup(n, 4242, edge_median=0.02 if n < 1000 else 0.004)
    dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
    b.refit(dm)
    tlo, thi = O.tri_aabb(moved)
    lo, hi = O.refit(ref["parent"], ref["left"], ref["right"], tlo, thi)
    ref_moved = dict(ref, lo=lo, hi=hi)
    got = b.export()
    assert np.array_equal(got["lo"].view(np.uint32), lo.view(np.uint32))
    orig, dirs = scenes.random_rays(6000, 8, origin=(0.0, 0.0, 0.0), spread=0.4)
    o_t, d_t = torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda()
    for any_hit in (False, True):
        tri = b.trace(dm, o_t, d_t, rtamd.RT_TRACE_ANY if any_hit else rtamd.RT_TRACE_CLOSEST)[0]
        torch.cuda.synchronize()
        rtri = O.trace(moved, ref_moved, orig, dirs, any_hit=any_hit)[0]
        assert np.array_equal(tri.cpu().numpy(), rtri), f"any={any_hit}: {np.sum(tri.cpu().numpy() != rtri)} rays"
        assert (rtri >= 0).sum() > 100
    assert b.diagnostics() == (0, 0)


def test_megakernel_after_refit_and_rebuild(ctx):
    """The megakernel's binary node records are filled on demand: the first
    megakernel render after a refit fills them from the refit boxes, and from
    then on every refit and rebuild keeps them current. Stale records would
    make its image differ from the wavefront's, which reads the quantized
    4-wide nodes."""
    _say("megakernel after refit / rebuild")
    hm = scenes.soup_mesh(30_000)
    mats, light, cam, opts = scenes.c2_setup(64, 48, 2)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 63)
    md = rtamd.materials_tensor(mats)
    for frame in range(4):
        moved = rtamd.gen_jitter(hm, 1e-3, 2000 + frame)
        dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
        if frame == 2:
            b.rebuild(dm)
        else:
            b.refit(dm)
        imgs = []
        for wf in (False, True):
            o = rtamd.render_opts(opts.spp, opts.depth, opts.seed,
                                  opts.flags | (rtamd.RT_RENDER_WAVEFRONT if wf else 0))
            imgs.append(rtamd.render(ctx, b, dm, md, len(mats), light, cam, o).cpu().numpy())
        assert np.array_equal(imgs[0], imgs[1]), f"frame {frame}: megakernel and wavefront images differ"
        assert imgs[0].max() > 0.1
    assert b.diagnostics() == (0, 0)


def test_captured_rebuild_then_megakernel_layout(ctx):
    """A rebuild captured in a hipGraph before any megakernel render: the
    binary records the megakernel reads are first requested after the
    capture, so the graph does not write them. Every later megakernel render
    of that BVH must still see the replayed geometry (the library refills
    the records before each such render), for two replays with moved
    vertices, against the oracle and the wavefront image."""
    _say("captured rebuild, then megakernel renders across replays")
    hm = scenes.soup_mesh(20_000, seed=17)
    mats, light, cam, opts = scenes.c2_setup(64, 48, 2)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, 63)
    md = rtamd.materials_tensor(mats)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            b.rebuild(dm)
    torch.cuda.current_stream().wait_stream(side)
    prev = None
    for frame in range(3):
        moved = rtamd.gen_jitter(hm, 0.02, 900 + frame)  # keeps the normals; moves triangles across boxes
        dm.set_positions(*[torch.from_numpy(a).cuda() for a in (moved.vx, moved.vy, moved.vz)])
        g.replay()
        torch.cuda.synchronize()
        mega = rtamd.render(ctx, b, dm, md, len(mats), light, cam, opts).cpu().numpy()
        o = rtamd.render_opts(opts.spp, opts.depth, opts.seed, opts.flags | rtamd.RT_RENDER_WAVEFRONT)
        wave = rtamd.render(ctx, b, dm, md, len(mats), light, cam, o).cpu().numpy()
        ref, _ = O.render(moved, O.build(moved, 63), mats, light, cam, opts)
        assert np.abs(mega - ref).max() <= 1e-4, f"frame {frame}: megakernel image is stale"
        assert np.array_equal(mega, wave), f"frame {frame}"
        assert prev is None or not np.array_equal(mega, prev)
        prev = mega
    assert b.diagnostics() == (0, 0)


def test_build_duplicate_keys(ctx):
    """Every triangle in a tiny region: 30-bit codes collide, the index fallback orders the tree."""
    _say("build duplicate keys")
    hm = scenes.duplicate_key_mesh(20_000)
    dm = rtamd.DeviceMesh(hm)
    for bits in (30, 63):
        b = rtamd.BVH(ctx, dm, bits)
        got = b.export()
        ref = O.build(hm, bits)
        assert len(np.unique(ref["codes"])) < hm.n
        _compare_build(got, ref, hm)


def test_build_flat_axis(ctx):
    """All triangles in the plane z = -1: (c - min)/0 is NaN and clamps to 0."""
    _say("build flat axis")
    hm = scenes.soup_mesh(5000)
    hm.vz[:] = -1.0
    dm = rtamd.DeviceMesh(hm)
    for bits in (30, 63):
        b = rtamd.BVH(ctx, dm, bits)
        _compare_build(b.export(), O.build(hm, bits), hm)


def _trace_compare(ctx, hm, bits, orig, dirs):
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, bits)
    ref_b = O.build(hm, bits)
    o_t = torch.from_numpy(orig).cuda()
    d_t = torch.from_numpy(dirs).cuda()
    for any_hit in (False, True):
        tri, t, a, be = b.trace(dm, o_t, d_t, rtamd.RT_TRACE_ANY if any_hit else rtamd.RT_TRACE_CLOSEST)
        torch.cuda.synchronize()
        tri, t, a, be = [x.cpu().numpy() for x in (tri, t, a, be)]
        rtri, rt_, ra, rb, _, _ = O.trace(hm, ref_b, orig, dirs, any_hit=any_hit)
        assert b.diagnostics() == (0, 0)
        assert np.array_equal(tri, rtri), f"hit ids differ (any={any_hit}): {np.sum(tri != rtri)} rays"
        h = rtri >= 0
        np.testing.assert_allclose(t[h], rt_[h], rtol=1e-5)
        np.testing.assert_allclose(a[h], ra[h], rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(be[h], rb[h], rtol=1e-5, atol=1e-12)
    return int(h.sum())


@pytest.mark.parametrize("bits", [30, 63])
def test_trace_parity_c1(ctx, bits):
    _say(f"trace parity c1 bits={bits}")
    hm = scenes.c1_mesh(32)
    orig, dirs = scenes.random_rays(20_000, 5)
    hits = _trace_compare(ctx, hm, bits, orig, dirs)
    assert hits > 1000


def test_trace_parity_soup(ctx):
    _say("trace parity soup")
    hm = scenes.soup_mesh(100_000)
    orig, dirs = scenes.random_rays(20_000, 6, spread=0.4)
    hits = _trace_compare(ctx, hm, 63, orig, dirs)
    assert hits > 1000


def test_trace_from_inside(ctx):
    """Origins inside the soup: negative-t hits count under the reference rule."""
    _say("trace from inside")
    hm = scenes.soup_mesh(50_000)
    rng = np.random.default_rng(9)
    orig = np.stack([rng.uniform(-.5, .5, 5000), rng.uniform(-.5, .5, 5000), rng.uniform(-2, -1, 5000)], 1)
    dirs = rng.normal(size=(5000, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    _trace_compare(ctx, hm, 63, orig, dirs)


def test_trace_far_origins(ctx):
    """Origins 10^3 .. 10^7 away from a small soup: the conservative fp32
    child test then widens every box by E ~ 6 * 2^-24 * |o| / |d|, more than
    the small nodes' extent, so the inverted boxes of unused wide-node slots
    can pass and their null leaf (a zero triangle) is tested; the hits must
    still equal the oracle's."""
    _say("trace far origins")
    hm = rtamd.gen_soup(3000, 13, edge_median=0.05)  # larger triangles: most rays hit one
    rng = np.random.default_rng(41)
    n = 4000
    dist = 10.0 ** rng.uniform(3, 7, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    target = np.column_stack([rng.uniform(-.5, .5, n), rng.uniform(-.5, .5, n), rng.uniform(-2, -1, n)])
    orig = target + u * dist[:, None]
    dirs = target - orig
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    hits = _trace_compare(ctx, hm, 63, orig, dirs)
    assert hits > 1000


@pytest.mark.parametrize("scale", [1.0, 5e4])
def test_trace_axis_aligned_and_tiny_directions(ctx, scale):
    """Directions with exactly zero, tiny (fp32-denormal, FP64-denormal) and
    huge-ratio components, from origins on the ground quad's plane, on box
    planes of the sphere and inside the scene: the reference's 1/d is then
    +-inf or enormous and its slab terms +-inf or NaN, which the traversal's
    fp32 child tests must never reject where the FP64 test accepts. Scaled
    by 5e4, the scene is wide enough that a ray with a zero direction
    component decodes the quantized planes by byte conversions while the
    other rays of its wave take the fp16 decode."""
    _say(f"trace axis-aligned and tiny directions scale={scale}")
    hm = scenes.c1_mesh(16)
    if scale != 1.0:
        hm = rtamd.HostMesh(hm.vx * scale, hm.vy * scale, hm.vz * scale, hm.nx, hm.ny, hm.nz, hm.idx, hm.mat_id)
    rng = np.random.default_rng(21)
    base = []
    for ax in range(3):
        for sgn in (1.0, -1.0):
            d = np.zeros(3)
            d[ax] = sgn
            base.append(d)
    for tiny in (1e-20, 1e-40, 1e-310, 0.0):
        for ax in range(3):
            d = rng.normal(size=3)
            d[ax] = tiny
            base.append(d / np.linalg.norm(d))
            d2 = np.zeros(3)
            d2[ax] = 1.0
            d2[(ax + 1) % 3] = tiny
            base.append(d2)
    base = np.array(base)
    origins = [np.array([0.0, 0.0, 0.0]), np.array([0.0, -1.0, -4.0]), np.array([0.25, -1.0, -2.0]),
               np.array([0.0, 0.0, -5.0]), np.array([1.0, 0.0, -4.0]), np.array([-4.0, -1.0, -8.0])]
    for _ in range(30):
        origins.append(np.array([rng.uniform(-3, 3), rng.choice([-1.0, rng.uniform(-1, 1)]), rng.uniform(-7, -1)]))
    orig = np.repeat(np.array(origins), len(base), axis=0) * scale
    dirs = np.tile(base, (len(origins), 1))
    hits = _trace_compare(ctx, hm, 30, orig, dirs)
    assert hits > 100


@pytest.mark.parametrize("hittable", [(54,), (20, 54)])
def test_trace_deep_chain_spills_stack(ctx, hittable):
    """A chain-shaped tree 56 levels deep whose every box the rays cross, with
    the hittable triangles at the far end of the traversal order: the stack
    grows past the 32 entries a lane keeps in LDS, moves halves to the spill
    area and brings them back, and the hits must still equal the oracle's."""
    _say(f"trace deep chain hittable={hittable}")
    hm, _ = scenes.deep_chain_mesh(hittable)
    parent = O.build(hm, 63)["parent"]
    depth = 0
    for p in range(hm.n):
        x, d = hm.n - 1 + p, 0
        while x != 0:
            x, d = parent[x], d + 1
        depth = max(depth, d)
    assert depth == hm.n - 1  # the keys make a chain
    rng = np.random.default_rng(31)
    xy = rng.uniform(0.45, 0.499, size=(64, 2))
    for z, dz in ((-1.0, 1.0), (2.0, -1.0)):
        orig = np.column_stack([xy, np.full(len(xy), z)])
        dirs = np.tile([1e-4, 1e-4, dz], (len(xy), 1))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        assert _trace_compare(ctx, hm, 63, orig, dirs) == len(xy)


def _render_compare(ctx, hm, bits, mats, light, cam, opts, pixels=None):
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH(ctx, dm, bits)
    md = rtamd.materials_tensor(mats)
    ref_b = O.build(hm, bits)
    ref, counts = O.render(hm, ref_b, mats, light, cam, opts)
    images = []
    for wavefront in (False, True):
        o = rtamd.render_opts(opts.spp, opts.depth, opts.seed,
                              opts.flags | (rtamd.RT_RENDER_WAVEFRONT if wavefront else 0))
        cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
        rgb = rtamd.render(ctx, b, dm, md, len(mats), light, cam, o, counters=cnt)
        torch.cuda.synchronize()
        got = rgb.cpu().numpy()
        diff = np.abs(got - ref)
        assert diff.max() <= 1e-4, \
            f"wavefront={wavefront} max pixel diff {diff.max()} at {np.unravel_index(diff.argmax(), diff.shape)}"
        c = cnt.cpu().numpy()
        assert c[0] == counts[0] and c[1] == counts[1] and c[2] == counts[2], (wavefront, c, counts)
        images.append(got)
    assert np.array_equal(images[0], images[1]), "megakernel and wavefront images differ"
    assert b.diagnostics() == (0, 0)
    return images[0], np.abs(images[0] - ref)


def test_render_parity_c1(ctx):
    _say("render parity c1 128x128 x2spp")
    hm = scenes.c1_mesh(32)
    mats, light, cam, opts = scenes.c1_setup(128, 128, 2)
    got, diff = _render_compare(ctx, hm, 30, mats, light, cam, opts)
    assert got.max() > 0.1


def test_render_parity_thinlens_soup(ctx):
    _say("render parity thin-lens soup 96x64 x3spp")
    hm = scenes.soup_mesh(60_000)
    mats, light, cam, opts = scenes.c2_setup(96, 64, 3)
    _render_compare(ctx, hm, 63, mats, light, cam, opts)


@pytest.mark.parametrize("where", ["centre", "inside", "near"])
def test_render_light_inside_scene_box(ctx, where):
    """Shadow rays are ordered by their direction from the light. With the
    light at the scene box's centre or inside it the direction cells are
    undefined (every key 0); just outside a face they span a wide angle.
    The order never changes a result, so the image must still equal the
    oracle's."""
    _say(f"render light {where}")
    hm = scenes.soup_mesh(20_000)
    mats, light, cam, opts = scenes.c2_setup(64, 48, 2)
    lo = np.array([hm.vx.min(), hm.vy.min(), hm.vz.min()], np.float64)
    hi = np.array([hm.vx.max(), hm.vy.max(), hm.vz.max()], np.float64)
    c = 0.5 * (lo + hi)
    pos = {"centre": c, "inside": lo + 0.3 * (hi - lo), "near": c + np.array([0.6 * (hi[0] - lo[0]), 0.0, 0.0])}
    light.position[:] = [float(x) for x in pos[where]]
    _render_compare(ctx, hm, 63, mats, light, cam, opts)


def test_render_parity_bounces(ctx):
    _say("render parity bounces depth 3")
    hm = scenes.displaced_mesh(20)
    mats, light, cam, opts = scenes.c1_setup(64, 64, 2)
    opts.depth = 3
    _render_compare(ctx, hm, 30, mats, light, cam, opts)
