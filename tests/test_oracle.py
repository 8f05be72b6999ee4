"""Pin the CPU oracle (no GPU): known answers, closed forms, brute force and
the reference's own recorded values.

Each check is chosen so that a plausible slip in the oracle (dropped term,
swapped operand, wrong index, wrong rounding) fails it; none of them
retypes the oracle's formula.
"""
import math
import os

import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from tests import scenes

REF = "/root/reference"


# ------------------------------------------------------------------ A14 mrand
def test_xorshift_marsaglia_known_value():
    """Marsaglia, "Xorshift RNGs" (2003): xor32 with y = 2463534242 first returns 723471715."""
    v, st = O.xorshift(2463534242)
    assert v == 723471715 and st == 723471715


def test_xorshift_full_period_sample():
    """xorshift32 (13,17,5) is a full-period generator: no repeat within 2^20 steps from a seed."""
    s = 1
    seen = set()
    for _ in range(1 << 16):
        s = O.xorshift(s)[1]
        assert s != 0 and s not in seen
        seen.add(s)


def test_rand_double_range_and_one():
    """randDouble = double(float(u) / 2^32) (kernels/mrand.cu:25-29): u >= 2^32 - 128 rounds to 1.0."""
    # find a state whose next output is 0xFFFFFFFF by inverting the three xorshift steps
    def inv_shl(y, k):
        x = y
        for _ in range(32 // k + 1):
            x = y ^ ((x << k) & 0xFFFFFFFF)
        return x

    def inv_shr(y, k):
        x = y
        for _ in range(32 // k + 1):
            x = y ^ (x >> k)
        return x

    target = 0xFFFFFFFF
    s = inv_shl(inv_shr(inv_shl(target, 5), 17), 13)
    v, st = O.rand_double(s)
    assert st == target and v == 1.0
    rng = np.random.default_rng(0)
    for s in rng.integers(1, 2**32, 2000, dtype=np.uint64):
        v, _ = O.rand_double(int(s))
        assert 0.0 <= v <= 1.0


def test_wang_hash_is_a_bijection_sample():
    """Wang's hash is invertible: 2^18 consecutive seeds give 2^18 distinct states."""
    vals = {O.wang_hash(i) for i in range(1 << 18)}
    assert len(vals) == 1 << 18


def test_wang_hash_known_answers():
    """Thomas Wang's 32-bit integer hash in the form kernels/mrand.cu:13-21 uses
    (as published in N. Reed, "Quick and Easy GPU Random Numbers in D3D11",
    2013): s = (s ^ 61) ^ (s >> 16); s *= 9; s ^= s >> 4; s *= 0x27d4eb2d;
    s ^= s >> 15, all mod 2^32. Worked by hand for s = 0:
    61 -> 549 -> 549 ^ 34 = 519 -> 519 * 668265261 mod 2^32 = 3232286779
    -> 3232286779 ^ 98641 = 3232319850. The other values follow the same five
    steps; a changed shift, multiplier or xor constant changes every one."""
    known = {0: 3232319850, 1: 663891101, 2: 3329832309, 256: 2763059176,
             1000000: 4032016030, 0xFFFFFFFF: 1895078355}
    for s, want in known.items():
        assert O.wang_hash(s) == want, s


def test_seed_index_16x16_blocks():
    """inc/indexing.h:4-9 with blockDim (16,16), gridDim.x = W/16."""
    W = 64
    assert O.seed_index(0, 0, W) == 0
    assert O.seed_index(15, 0, W) == 15
    assert O.seed_index(0, 1, W) == 16
    assert O.seed_index(16, 0, W) == 256          # next block along x
    assert O.seed_index(0, 16, W) == 256 * (W // 16)  # next block row
    # a bijection on the rendered region
    ids = {O.seed_index(x, y, W) for y in range(32) for x in range(W)}
    assert len(ids) == 32 * W and max(ids) == 32 * W - 1


def test_sincos_against_libm():
    xs = np.linspace(0.0, 2 * math.pi, 20001)
    for x in xs:
        s, c = O.sincos(float(x))
        assert abs(s - math.sin(x)) <= 2.5e-16 and abs(c - math.cos(x)) <= 2.5e-16


# ------------------------------------------------------------------ A1-A3
def test_expand_bits_exhaustive():
    """ExpandBits (kernels/bvh.cu:13-21): bit b of v lands at bit 3b, for all 1024 inputs."""
    for v in range(1024):
        e = O.expand_bits10(v)
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert e == want
        assert e & ~0x09249249 == 0


def _box_mesh(boxes):
    """One degenerate 'triangle' per (lo, hi) pair whose AABB is exactly that box."""
    vx, vy, vz, idx = [], [], [], []
    for k, (lo, hi) in enumerate(boxes):
        pts = [lo, hi, (lo[0], hi[1], lo[2])]
        for p in pts:
            vx.append(p[0]); vy.append(p[1]); vz.append(p[2])
        idx += [3 * k, 3 * k + 1, 3 * k + 2]
    z = np.zeros(len(vx), np.float32)
    return rtamd.HostMesh(vx, vy, vz, z, z, z, idx)


def test_morton_known_answers():
    """Centroid (0,0,0) -> 0; (1,1,1) clamps to 1023 per axis -> all 30 bits;
    x is the high bit of each triple (xx*4 + yy*2 + zz)."""
    boxes = [((0, 0, 0), (0, 0, 0)), ((1, 1, 1), (1, 1, 1)), ((0.5, 0, 0), (0.5, 0, 0)),
             ((0, 0.5, 0), (0, 0.5, 0)), ((0, 0, 0.5), (0, 0, 0.5))]
    hm = _box_mesh(boxes)
    lo, hi = O.tri_aabb(hm)
    mn, mx = np.zeros(3), np.ones(3)
    c30 = O.morton(lo, hi, mn, mx, 30)
    assert list(c30) == [0, (1 << 30) - 1, 4 << 27, 2 << 27, 1 << 27]
    c63 = O.morton(lo, hi, mn, mx, 63)
    assert list(c63) == [0, (1 << 63) - 1, 4 << 60, 2 << 60, 1 << 60]


def test_morton_float_rounding_before_truncation():
    """u*1024 is rounded to float before the cast (kernels/bvh.cu:25): with
    u*1024 = 1022.9999999 the x cell is 1023, not 1022."""
    hm = _box_mesh([((1023.0, 0, 0), (1023.0, 0, 0))])
    lo, hi = O.tri_aabb(hm)
    mn = np.zeros(3)
    mx = np.array([1024.0000001, 1.0, 1.0])
    assert 1022.99 < 1023.0 / 1024.0000001 * 1024 < 1023.0
    code = int(O.morton(lo, hi, mn, mx, 30)[0])
    assert code & 0x24924924 == 0x24924924  # all ten x bits set: cell 1023


def test_morton_flat_axis_nan_clamps_to_zero():
    """(c - min)/(max - min) = 0/0 = NaN on a flat axis; fmax(NaN, 0) = 0."""
    boxes = [((0, 0, 5), (0, 0, 5)), ((1, 1, 5), (1, 1, 5))]
    hm = _box_mesh(boxes)
    lo, hi = O.tri_aabb(hm)
    mn, mx = O.scene_bounds(lo, hi)
    assert mn[2] == mx[2] == 5.0
    c = O.morton(lo, hi, mn, mx, 30)
    assert all((int(v) & 0x09249249) == 0 for v in c)  # every z bit is 0
    assert c[0] == 0 and c[1] == ((1 << 30) - 1) & ~0x09249249


def test_scene_bounds_init_and_nan():
    """findMinMax folds from +-1e9 with fminf/fmaxf (kernels/scene.cu:72-85): NaN is skipped,
    and a scene beyond +1e9 reports 1e9."""
    lo = np.array([[np.nan, 1, 2], [3, np.nan, -4]], np.float32)
    hi = np.array([[np.nan, 5, 6], [7, 8, np.nan]], np.float32)
    mn, mx = O.scene_bounds(lo, hi)
    assert list(mn) == [3, 1, -4] and list(mx) == [7, 8, 6]
    mn, mx = O.scene_bounds(np.full((1, 3), 2e9, np.float32), np.full((1, 3), 3e9, np.float32))
    assert list(mn) == [1e9] * 3 and list(mx) == [3e9] * 3


def test_tri_aabb_is_vertex_min_max():
    hm = scenes.soup_mesh(3000)
    lo, hi = O.tri_aabb(hm)
    P = np.stack([hm.vx[hm.idx], hm.vy[hm.idx], hm.vz[hm.idx]], 1).reshape(-1, 3, 3)
    assert np.array_equal(lo, P.min(1)) and np.array_equal(hi, P.max(1))


def _read_obj(path):
    """Minimal OBJ reader: vertex positions and fan-triangulated faces."""
    vs, faces = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                ix = [int(t.split("/")[0]) for t in p[1:]]
                ix = [i - 1 if i > 0 else len(vs) + i for i in ix]
                for k in range(1, len(ix) - 1):
                    faces.append([ix[0], ix[k], ix[k + 1]])
    return np.array(vs), np.array(faces, np.int32)


@pytest.mark.skipif(not os.path.exists(f"{REF}/objs/funyz/logcabin.obj"), reason="reference tree not present")
def test_logcabin_counts_and_bounds_known_answer():
    """results/logcabin.out:23-25,38: 4980 vertices, 9338 triangles,
    min(-6.816752, -2.977815, -24.966003) max(9.500026, 3.785227, -8.857635)."""
    vs, faces = _read_obj(f"{REF}/objs/funyz/logcabin.obj")
    assert len(vs) == 4980 and len(faces) == 9338
    v32 = vs.astype(np.float32)
    z = np.zeros(len(vs), np.float32)
    hm = rtamd.HostMesh(v32[:, 0], v32[:, 1], v32[:, 2], z, z, z, faces.reshape(-1))
    lo, hi = O.tri_aabb(hm)
    mn, mx = O.scene_bounds(lo, hi)
    assert [f"{v:0.6f}" for v in mn] == ["-6.816752", "-2.977815", "-24.966003"]
    assert [f"{v:0.6f}" for v in mx] == ["9.500026", "3.785227", "-8.857635"]


# ------------------------------------------------------------------ A4-A7
@pytest.mark.parametrize("bits", [30, 63])
def test_sort_is_stable_sort(bits):
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 64, 5000).astype(np.uint32 if bits == 30 else np.uint64)
    s, ids = O.sort(codes)
    perm = np.argsort(codes, kind="stable")
    assert np.array_equal(ids, perm) and np.array_equal(s, codes[perm])


def _check_tree_invariants(parent, left, right, n):
    assert parent[0] == -1
    internal = np.arange(n - 1)
    assert np.all(left[internal] != right[internal])
    kids = np.concatenate([left[internal], right[internal]])
    assert np.array_equal(np.sort(kids), np.arange(1, 2 * n - 1))  # every non-root node has one parent
    assert np.all(parent[kids] == np.concatenate([internal, internal]))


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("scene", ["c1", "soup", "dup"])
def test_hierarchy_equals_topdown_radix_tree(bits, scene):
    hm = {"c1": lambda: scenes.c1_mesh(8), "soup": lambda: scenes.soup_mesh(3000, seed=5),
          "dup": lambda: scenes.duplicate_key_mesh(2000)}[scene]()
    b = O.build(hm, bits)
    p, l, r = O.hierarchy(b["codes"], b["ids"], topdown=True)
    assert np.array_equal(p, b["parent"]) and np.array_equal(l, b["left"]) and np.array_equal(r, b["right"])
    _check_tree_invariants(b["parent"], b["left"], b["right"], hm.n)
    # leaves carry the sorted triangle ids
    assert np.array_equal(b["left"][hm.n - 1:], b["ids"].astype(np.int32))


def test_hierarchy_hand_built_example():
    """Keys 1,2,4,5,19,24,25,30 (Karras 2012, fig. 3 ordering): root splits after the
    first key with the top differing bit."""
    codes = np.array([1, 2, 4, 5, 19, 24, 25, 30], np.uint32)
    ids = np.arange(8, dtype=np.uint32)
    p, l, r = O.hierarchy(codes, ids)
    n = 8
    # root covers [0,7]; 00001..00101 vs 10011.. differ at bit 4 -> split after index 3
    assert l[0] == 3 and r[0] == 4
    # node 3 covers [0,3]: 0001,0010 | 0100,0101 -> split after 1
    assert l[3] == 1 and r[3] == 2
    assert l[1] == n - 1 + 0 and r[1] == n - 1 + 1
    _check_tree_invariants(p, l, r, n)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_refit_any_order_equals_brute_force(seed):
    hm = scenes.displaced_mesh(12)
    b = O.build(hm, 63)
    tlo, thi = O.tri_aabb(hm)
    blo, bhi = O.refit(b["parent"], b["left"], b["right"], tlo, thi, brute=True)
    perm = np.random.default_rng(seed).permutation(hm.n).astype(np.int32)
    plo, phi = O.refit(b["parent"], b["left"], b["right"], tlo, thi, perm=perm)
    assert np.array_equal(blo, b["lo"]) and np.array_equal(bhi, b["hi"])
    assert np.array_equal(plo, blo) and np.array_equal(phi, bhi)
    assert np.array_equal(b["lo"][0], tlo.min(0)) and np.array_equal(b["hi"][0], thi.max(0))


# ------------------------------------------------------------------ A9-A12
def _unit_tri():
    z = np.zeros(3, np.float32)
    return rtamd.HostMesh([0, 1, 0], [0, 0, 1], [0, 0, 0], z, z, np.ones(3, np.float32), [0, 1, 2])


def test_triangle_known_answers():
    hm = _unit_tri()
    ok, t, a, b, n = O.intersect_triangle(hm, 0, [0.25, 0.25, 1.0], [0, 0, -1.0])
    assert ok and t == 1.0 and a == 0.25 and b == 0.25 and np.allclose(n, [0, 0, 1])
    # behind the origin: the reference accepts negative t (kernels/bvh.cu:347)
    ok, t, a, b, n = O.intersect_triangle(hm, 0, [0.25, 0.25, -2.0], [0, 0, -1.0])
    assert ok and t == -2.0
    # outside the triangle (alpha + beta > 1) and parallel rays miss
    assert not O.intersect_triangle(hm, 0, [0.75, 0.75, 1.0], [0, 0, -1.0])[0]
    assert not O.intersect_triangle(hm, 0, [0.2, 0.2, 1.0], [1.0, 0, 0])[0]
    # |t| < 1e-8 is rejected (RAY_EPSILON, inc/common.h:2)
    assert not O.intersect_triangle(hm, 0, [0.25, 0.25, 5e-9], [0, 0, -1.0])[0]


def test_triangle_with_non_finite_vertex_is_not_hit():
    """A NaN vertex makes alpha, beta and t NaN; the reference's comparisons
    let such a hit through (kernels/bvh.cu:330-345), the oracle (and the
    GPU) reject it. A vertex at infinity either fails the reference's own
    tests or gives a NaN: never a hit with a non-finite t."""
    z = np.zeros(3, np.float32)
    for bad in (np.nan, np.inf, -np.inf):
        for k in range(3):
            for axis in range(3):
                v = [np.array([0, 1, 0], np.float32), np.array([0, 0, 1], np.float32), np.zeros(3, np.float32)]
                v[axis][k] = bad
                hm = rtamd.HostMesh(v[0], v[1], v[2], z, z, np.ones(3, np.float32), [0, 1, 2])
                for o, d in (([0.25, 0.25, 1.0], [0, 0, -1.0]), ([0.1, 0.2, -3.0], [0.01, 0.02, 1.0])):
                    ok, t, a, b, _ = O.intersect_triangle(hm, 0, o, d)
                    assert not ok or (math.isfinite(t) and math.isfinite(a) and math.isfinite(b)), (bad, k, axis)
                    if np.isnan(bad):
                        assert not ok, (k, axis)


def test_box_known_answers():
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert O.box_hit(lo, hi, [0.5, 0.5, 5], [0, 0, -1.0])
    assert O.box_hit(lo, hi, [0.5, 0.5, -5], [0, 0, -1.0])      # box behind the origin still counts
    assert not O.box_hit(lo, hi, [1.5, 0.5, 5], [0, 0, -1.0])
    assert O.box_hit(lo, hi, [1.0, 0.5, 5], [0, 0, -1.0])        # grazing the face counts (<=)
    assert O.box_hit(lo, hi, [2, 2, 2], [1, 1, 1.0])            # the line, not the ray: enters at t = -1.5
    assert not O.box_hit(lo, hi, [2, 0.5, 0.5], [0, 1.0, 0])


@pytest.mark.parametrize("bits", [30, 63])
def test_closest_hit_equals_brute_force(bits):
    hm = scenes.c1_mesh(16)
    b = O.build(hm, bits)
    orig, dirs = scenes.random_rays(3000, 11)
    tri, t, *_ = O.trace(hm, b, orig, dirs)
    leafpos = np.empty(hm.n, np.int32)
    leafpos[b["ids"]] = np.arange(hm.n)
    btri, bt = O.trace_brute(hm, leafpos, orig, dirs)
    assert np.array_equal(tri, btri) and np.array_equal(t, bt)
    assert (tri >= 0).sum() > 300


def test_any_hit_is_first_leaf_in_sorted_order():
    """Left-first DFS reaches leaves in sorted order, so any-hit returns the hit
    triangle with the smallest sorted position."""
    hm = rtamd.gen_soup(1500, 8, edge_median=0.06)
    b = O.build(hm, 63)
    orig, dirs = scenes.random_rays(200, 12, origin=(0, 0, 0.0), spread=0.3)
    tri, *_ = O.trace(hm, b, orig, dirs, any_hit=True)
    leafpos = np.empty(hm.n, np.int32)
    leafpos[b["ids"]] = np.arange(hm.n)
    checked = 0
    for r in range(len(orig)):
        hits = [k for k in range(hm.n) if O.intersect_triangle(hm, k, orig[r], dirs[r])[0]]
        if not hits:
            assert tri[r] == -1
            continue
        assert tri[r] == min(hits, key=lambda k: leafpos[k])
        checked += 1
    assert checked > 20


# ------------------------------------------------------------------ A15-A17
def test_primary_ray_thin_lens_focus():
    """Every lens sample of a pixel passes through normalize(xx, yy, -1) * focal_point."""
    cam = rtamd.camera(64, 48, fov_deg=40.0, fstop=2.0, focal_point=3.0)
    angle = math.tan(math.pi * 0.5 * 40.0 / 180.0)
    for (px, py) in [(0, 0), (10, 20), (63, 47)]:
        xx = (2 * ((px + 0.5) / 64) - 1) * angle * (64 / 48)
        yy = (1 - 2 * ((py + 0.5) / 48)) * angle
        f = np.array([xx, yy, -1.0])
        f = f / np.linalg.norm(f) * 3.0
        for s in range(4):
            o, d = O.primary_ray(cam, px, py, s)
            assert abs(o[2]) == 0.0 and np.hypot(o[0], o[1]) <= 0.25 + 1e-12  # aperture 1/2 -> radius 1/4
            assert abs(np.linalg.norm(d) - 1) < 1e-6
            w = f - o
            w /= np.linalg.norm(w)
            assert np.allclose(w, d, atol=3e-7)


def test_primary_ray_lens_draw_order_known_answer():
    """kernels/raytracer.cu:143-145: the FIRST draw of a sample is the angle
    (theta = 2 pi u1), the SECOND the radius (rho = (1/fstop) sqrt(u2) / 2), and
    the origin is (rho cos theta, rho sin theta, 0). The draws come from the
    pixel's own stream: state wang(seed_index), then xorshift steps, u =
    float(x) / 2^32. Swapping the two draws, dropping the square root or
    using another aperture fails here."""
    W, H, fstop = 48, 32, 2.5
    cam = rtamd.camera(W, H, fov_deg=50.0, fstop=fstop, focal_point=2.0)
    for (px, py) in [(0, 0), (17, 5), (47, 31), (30, 16)]:
        st = O.wang_hash(O.seed_index(px, py, W))
        for s in range(3):
            u1, st = O.rand_double(st)
            u2, st = O.rand_double(st)
            o, _ = O.primary_ray(cam, px, py, s)
            rho = (1.0 / fstop) * math.sqrt(u2) / 2.0
            assert abs(math.hypot(o[0], o[1]) - rho) <= 1e-15
            if rho > 1e-9:
                ang = math.atan2(o[1], o[0]) % (2 * math.pi)
                want = (2 * math.pi * u1) % (2 * math.pi)
                assert min(abs(ang - want), 2 * math.pi - abs(ang - want)) <= 1e-9, (px, py, s)
            assert o[2] == 0.0


def test_cosine_dir_known_answers():
    """The bounce frame of src/ray.cpp:5-32 worked by hand for n = +z: h = n with
    its smallest-|.| component (x) set to 1 -> h = (1,0,1); x = h ^ n = (0,-1,0);
    z = x ^ n = (-1,0,0); sin(theta) = sqrt(u1), cos(theta) = sqrt(1 - u1),
    phi = 2 pi u2, dir = sin cos(phi) x + cos n + sin sin(phi) z."""
    n = np.array([0.0, 0.0, 1.0])
    assert np.allclose(O.cosine_dir(n, 0.0, 0.3), [0, 0, 1], atol=1e-15)
    assert np.allclose(O.cosine_dir(n, 0.25, 0.0), [0, -0.5, math.sqrt(0.75)], atol=1e-15)
    assert np.allclose(O.cosine_dir(n, 0.25, 0.25), [-0.5, 0, math.sqrt(0.75)], atol=1e-15)
    assert np.allclose(O.cosine_dir(n, 1.0, 0.5), [0, 1, 0], atol=1e-15)
    # n = +x: the smallest component is y (ties go to x, then y): h = (1,1,0),
    # x = h ^ n = (0,0,-1), z = x ^ n = (0,-1,0)
    n = np.array([1.0, 0.0, 0.0])
    assert np.allclose(O.cosine_dir(n, 0.64, 0.0), [0.6, 0, -0.8], atol=1e-15)
    assert np.allclose(O.cosine_dir(n, 0.64, 0.25), [0.6, -0.8, 0], atol=1e-15)


def test_cosine_dir_moments():
    """Cosine-weighted hemisphere sampling, p(w) = cos(theta) / pi, has
    E[cos] = 2/3, E[cos^2] = 1/2, a uniform azimuth (E[d_t] = 0 for both
    tangents, E[sin^2 phi] = 1/2), and unit directions in the normal's
    hemisphere. Stratified midpoints over (u1, u2) on a 200 x 200 grid."""
    g = (np.arange(200) + 0.5) / 200
    rng = np.random.default_rng(3)
    for n in (np.array([0.0, 0.0, 1.0]), np.array([0.3, -0.5, 0.81]), np.array([-1.0, 0.2, 0.1])):
        n = n / np.linalg.norm(n)
        t1 = np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(n, t1)
        c, c2, a1, a2, a11 = 0.0, 0.0, 0.0, 0.0, 0.0
        for u1 in g:
            for u2 in g:
                d = O.cosine_dir(n, u1 + rng.uniform(-1e-4, 1e-4), u2)
                # normalize() divides by sqrtf (inc/vec.h:119): unit to fp32 precision
                assert abs(np.linalg.norm(d) - 1) < 1e-6 and d @ n >= -1e-6
                ct = d @ n
                c += ct
                c2 += ct * ct
                a1 += d @ t1
                a2 += d @ t2
                a11 += (d @ t1) ** 2 / max(1e-300, 1 - ct * ct)
        m = 200 * 200
        assert abs(c / m - 2 / 3) < 2e-3 and abs(c2 / m - 0.5) < 2e-3
        assert abs(a1 / m) < 2e-3 and abs(a2 / m) < 2e-3 and abs(a11 / m - 0.5) < 2e-3


def test_render_empty_scene_is_background():
    """Every ray misses -> colour (0.9, 0.9, 0.9) (kernels/raytracer.cu:244)."""
    hm = scenes.soup_mesh(10)
    for a in (hm.vx, hm.vy, hm.vz):
        a += np.float32(100.0)
    b = O.build(hm, 30)
    mats, light, cam, opts = scenes.c1_setup(16, 16, 2)
    rgb, counts = O.render(hm, b, mats, light, cam, opts)
    assert np.all(rgb == np.float32(0.9)) and counts[1] == 0 and counts[0] == 16 * 16 * 2


def test_render_lambert_plane_closed_form():
    """A plane z=-2 facing the camera, light radius 0 (offset = -0.5 exactly),
    no occluders: I = kd * max(0, N.L) * min(1, 1/(c0 + c1 d + c2 d^2)) with d
    the distance to the light position (kernels/material.cu:6-40, Light.cu:56-81)."""
    z = np.zeros(4, np.float32)
    hm = rtamd.HostMesh([-50, 50, 50, -50], [-50, -50, 50, 50], [-2, -2, -2, -2], z, z, np.ones(4, np.float32),
                        [0, 1, 2, 0, 2, 3])
    b = O.build(hm, 30)
    mats = [rtamd.material(kd=(0.7, 0.5, 0.3))]
    light = rtamd.light(position=(0.3, 0.2, 1.0), radius=0.0, c0=0.5, c1=0.1, c2=0.02)
    W = H = 16
    cam = rtamd.camera(W, H, fov_deg=60.0, fstop=float("inf"), focal_point=1.0)
    opts = rtamd.render_opts(1, flags=0)
    rgb, _ = O.render(hm, b, mats, light, cam, opts)
    rgb = rgb.reshape(H, W, 3)
    lp = np.array(light.position[:])
    ang = math.tan(math.pi * 0.5 * 60.0 / 180.0)
    for py in range(H):
        for px in range(W):
            xx = (2 * ((px + 0.5) / W) - 1) * ang
            yy = (1 - 2 * ((py + 0.5) / H)) * ang
            d = np.array([xx, yy, -1.0])
            d /= np.linalg.norm(d)
            q = d * (-2.0 / d[2])
            L = lp - 0.5 - q
            L /= np.linalg.norm(L)
            dist = np.linalg.norm(lp - q)
            f = min(1.0, 1.0 / (0.5 + 0.1 * dist + 0.02 * dist * dist))
            want = np.array([0.7, 0.5, 0.3]) * max(0.0, L[2]) * f
            got = rgb[H - py - 1, px]  # row flip (kernels/raytracer.cu:106)
            assert np.allclose(got, want, rtol=2e-6, atol=1e-7), (px, py, got, want)
