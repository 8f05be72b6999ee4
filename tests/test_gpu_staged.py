"""The staged build (rt_bvh_create, rt_bvh_compute_bounds,
rt_bvh_compute_morton, rt_bvh_sort_morton, rt_bvh_build_tree): the calls of
the reference's BVH_d::setUp one at a time (src/bvh.cpp:3-32,
kernels/bvh.cu:260-292), including computeMortonCodes with caller-given
bounds, against the oracle's stages; stage-order errors; and a C++ caller
shaped like the reference's Scene_d / BVH_d code, compiled against
include/rt_facade.hpp."""
import os
import subprocess

import numpy as np
import pytest

import rtamd
from oracle import oracle as O
from tests import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE_CALLER = os.path.join(ROOT, "tests", "native", "facade_caller")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


def _oracle_stages(hm, bits, mn=None, mx=None):
    """The reference's build from its stages, with the Morton bounds given or
    the scene bounds."""
    tlo, thi = O.tri_aabb(hm)
    smn, smx = O.scene_bounds(tlo, thi)
    if mn is None:
        mn, mx = smn, smx
    codes = O.morton(tlo, thi, mn, mx, bits)
    codes, ids = O.sort(codes)
    parent, left, right = O.hierarchy(codes, ids)
    lo, hi = O.refit(parent, left, right, tlo, thi)
    return dict(codes=codes, ids=ids, parent=parent, left=left, right=right, lo=lo, hi=hi, tri_lo=tlo, tri_hi=thi,
                mn=smn, mx=smx)


def _assert_export_equal(got, want, what):
    for k in ("codes", "ids", "parent", "left", "right"):
        g = got[k].astype(np.int64) if k != "codes" else got[k].view(want[k].dtype)
        assert np.array_equal(g.reshape(-1), want[k].astype(g.dtype).reshape(-1)), f"{what}: {k} differs"
    for k in ("lo", "hi"):
        assert np.array_equal(got[k].reshape(-1).view(np.uint32), want[k].reshape(-1).view(np.uint32)), \
            f"{what}: {k} differs"


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("n", [1, 2, 513, 40_000])
def test_staged_equals_fused_and_oracle(ctx, bits, n):
    _say(f"staged build n={n} bits={bits}: scene bounds, then caller bounds")
    hm = scenes.soup_mesh(n, seed=100 + n)
    dm = rtamd.DeviceMesh(hm)
    fused = rtamd.BVH(ctx, dm, bits).export()
    b = rtamd.BVH.create(ctx, n, bits)
    tlo = torch.empty(n * 3, dtype=torch.float32, device="cuda")
    thi = torch.empty(n * 3, dtype=torch.float32, device="cuda")
    b.compute_bounds(dm, tlo, thi)
    mn, mx = b.scene_bounds()
    want = _oracle_stages(hm, bits)
    assert np.array_equal(mn, want["mn"]) and np.array_equal(mx, want["mx"])
    assert np.array_equal(tlo.cpu().numpy().view(np.uint32), want["tri_lo"].reshape(-1).view(np.uint32))
    assert np.array_equal(thi.cpu().numpy().view(np.uint32), want["tri_hi"].reshape(-1).view(np.uint32))
    b.compute_morton()
    b.sort_morton()
    b.build_tree(dm)
    staged = b.export()
    for k in fused:
        assert np.array_equal(staged[k], fused[k]), f"staged {k} differs from rt_bvh_build"
    _assert_export_equal(staged, want, "scene bounds")
    # computeMortonCodes(mMin, mMax) with bounds the caller chose: wider than
    # the scene and not representable in fp32
    cmn, cmx = mn - np.array([0.1, 0.3, 1.0 / 3.0]), mx + np.array([0.7, 0.01, 2.0 / 7.0])
    b.compute_bounds(dm)
    b.compute_morton(cmn, cmx)
    b.sort_morton()
    b.build_tree(dm)
    _assert_export_equal(b.export(), _oracle_stages(hm, bits, cmn, cmx), "caller bounds")
    assert b.diagnostics() == (0, 0)
    # the staged tree traces like the oracle's tree of the same keys
    ref = _oracle_stages(hm, bits, cmn, cmx)
    orig, dirs = scenes.random_rays(3000, 3, origin=(0.0, 0.0, 0.0), spread=0.5)
    tri = b.trace(dm, torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda(), rtamd.RT_TRACE_CLOSEST)[0]
    rtri = O.trace(hm, ref, orig, dirs)[0]
    assert np.array_equal(tri.cpu().numpy(), rtri)


def test_stage_order_is_enforced(ctx):
    _say("staged build: out-of-order calls return RT_ERR_INVALID")
    hm = scenes.soup_mesh(1000, seed=7)
    dm = rtamd.DeviceMesh(hm)
    b = rtamd.BVH.create(ctx, hm.n, 30)

    def invalid(fn):
        with pytest.raises(rtamd.RtError) as e:
            fn()
        assert e.value.status == rtamd.RT_ERR_INVALID, e.value

    orig, dirs = scenes.random_rays(10, 1)
    o_t, d_t = torch.from_numpy(orig).cuda(), torch.from_numpy(dirs).cuda()
    invalid(lambda: b.scene_bounds())
    invalid(lambda: b.compute_morton())
    invalid(lambda: b.sort_morton())
    invalid(lambda: b.build_tree(dm))
    invalid(lambda: b.trace(dm, o_t, d_t, rtamd.RT_TRACE_CLOSEST))
    invalid(lambda: b.refit(dm))
    invalid(lambda: b.export())
    invalid(lambda: b.compute_morton([0, 0, 0], None))
    b.compute_bounds(dm)
    invalid(lambda: b.sort_morton())
    b.compute_morton()
    invalid(lambda: b.build_tree(dm))
    b.sort_morton()
    invalid(lambda: b.sort_morton())
    b.build_tree(dm)
    invalid(lambda: b.build_tree(dm))
    b.refit(dm)
    # a new Morton stage leaves the BVH unbuilt until its tree is built
    b.compute_morton()
    invalid(lambda: b.trace(dm, o_t, d_t, rtamd.RT_TRACE_CLOSEST))
    b.rebuild(dm)
    tri = b.trace(dm, o_t, d_t, rtamd.RT_TRACE_CLOSEST)[0]
    torch.cuda.synchronize()
    assert np.array_equal(tri.cpu().numpy(), O.trace(hm, O.build(hm, 30), orig, dirs)[0])
    for bad in (0, -1, 1 << 30):
        with pytest.raises(rtamd.RtError):
            rtamd.BVH.create(ctx, bad, 30)
    with pytest.raises(rtamd.RtError):
        rtamd.BVH.create(ctx, 10, 31)


def test_reference_shaped_cpp_caller(tmp_path):
    """tests/native/facade_caller.cpp drives the build the way the
    reference's Scene_d::operator= and BVH_d::setUp do; it checks on its own
    that three routes give the same tree and writes the tree it built with
    caller-chosen bounds, which must equal the oracle's."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    if not os.path.exists(FACADE_CALLER):
        pytest.fail(f"{FACADE_CALLER} is not built (run __graft_entry__.build())")
    n, seed = 30_000, 4242
    _say(f"C++ reference-shaped caller n={n}")
    out = subprocess.run([FACADE_CALLER, str(n), str(seed), str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "routes agree" in out.stdout, out.stdout
    hm = rtamd.gen_soup(n, seed)
    raw = lambda name, dt: np.fromfile(os.path.join(tmp_path, name), dtype=dt)
    mn, mx = raw("bounds.bin", np.float64).reshape(2, 3)
    want = _oracle_stages(hm, 30, mn, mx)
    got = dict(codes=raw("codes.bin", np.uint32), ids=raw("ids.bin", np.uint32), parent=raw("parent.bin", np.int32),
               left=raw("left.bin", np.int32), right=raw("right.bin", np.int32), lo=raw("lo.bin", np.float32),
               hi=raw("hi.bin", np.float32))
    _assert_export_equal(got, want, "C++ caller, caller bounds")
    smn, smx = raw("scene_bounds.bin", np.float64).reshape(2, 3)
    assert np.array_equal(smn, want["mn"]) and np.array_equal(smx, want["mx"])
    # computeBoundingBoxes' BBoxs, in triangle order
    assert np.array_equal(raw("bbox_lo.bin", np.uint32), want["tri_lo"].reshape(-1).view(np.uint32))
    assert np.array_equal(raw("bbox_hi.bin", np.uint32), want["tri_hi"].reshape(-1).view(np.uint32))
