"""The fp32 inner leaf-box test of the persistent traversal (inner_slab in
cuda-rt_amd/csrc/rt_trace.hip) may stand in for the reference's FP64 slab
test (inc/bbox.h:82-102) only when it implies it: a box the inner test
accepts must be accepted by the FP64 test. Checked on the CPU against the
oracle's FP64 box test over random, grazing, axis-aligned, tiny-direction,
far-origin and flat-box cases, with the kernel's fp32 ray terms
(make_rayf / rayf_axis) emulated in numpy and each fused multiply-add
widened by one fp32 ulp in the direction that accepts more boxes, so the
emulated test accepts everything the kernel's does.
"""
import numpy as np

from oracle import oracle as O

EPS24 = 5.9604644775390625e-08


def rayf(o, d, bmax):
    """fp32 1/d and the near/far offsets o/d +- E of one ray (rayf_axis)."""
    iv, nn, ff, pos = [], [], [], []
    for k in range(3):
        with np.errstate(divide="ignore"):
            inv = np.float64(1.0) / np.float64(d[k])
        pos.append(bool(inv >= 0))
        if not abs(inv) <= 1e30:
            inv = np.copysign(1e30, d[k])
        oi = o[k] * inv
        E = 6.0 * EPS24 * (bmax + abs(o[k])) * abs(inv)
        iv.append(np.float32(inv))
        nn.append(np.float32(oi + E))
        ff.append(np.float32(oi - E))
    return iv, nn, ff, pos


def fma_lo(a, b, c):
    v = np.float32(np.float64(a) * np.float64(b) + np.float64(c))
    return np.nextafter(v, np.float32(-np.inf))


def fma_hi(a, b, c):
    v = np.float32(np.float64(a) * np.float64(b) + np.float64(c))
    return np.nextafter(v, np.float32(np.inf))


def inner_accepts(lo, hi, o, d, bmax):
    iv, nn, ff, pos = rayf(o, d, bmax)
    t0 = -np.inf
    t1 = np.inf
    for k in range(3):
        near, far = (lo[k], hi[k]) if pos[k] else (hi[k], lo[k])
        t0 = max(t0, fma_lo(near, iv[k], -ff[k]))
        t1 = min(t1, fma_hi(far, iv[k], -nn[k]))
    return t0 <= t1


def _cases(rng, n):
    for _ in range(n):
        kind = rng.integers(0, 8)
        c = rng.uniform(-1.5, 1.5, 3)
        size = 10.0 ** rng.uniform(-4, -1)
        v = (c + rng.normal(0, size, (3, 3))).astype(np.float32)
        if kind == 5:
            v[:, rng.integers(0, 3)] = np.float32(c[0])  # flat box
        lo, hi = v.min(axis=0), v.max(axis=0)
        bmax = float(max(np.abs(lo).max(), np.abs(hi).max(), 1.5))
        o = rng.uniform(-3, 3, 3)
        if kind == 1:
            o = o * 300.0  # far origin
        if kind == 2:
            o = (lo + (hi - lo) * rng.uniform(0, 1, 3)).astype(np.float64)  # inside the box
        target = lo + (hi - lo) * rng.uniform(-0.05, 1.05, 3)  # grazing the boundary at times
        if kind == 6:
            # aimed at a corner, an edge or a face to within a few ulps
            corner = np.where(rng.integers(0, 2, 3) == 1, hi, lo).astype(np.float64)
            keep = rng.integers(0, 2, 3) == 1
            target = np.where(keep, target, corner) * (1 + rng.normal(0, 1e-7, 3))
        if kind == 7:
            o[rng.integers(0, 3)] = float(lo[0]) if rng.integers(0, 2) else float(hi[0])  # origin on a plane
        d = np.asarray(target, np.float64) - o
        if kind == 3:
            d[rng.integers(0, 3)] = 0.0
        if kind == 4:
            d[rng.integers(0, 3)] = rng.choice([1e-35, -1e-35, 5e-324, -5e-324, 1e-12])
        nrm = np.linalg.norm(d)
        if nrm == 0:
            continue
        yield lo, hi, o, d / nrm, bmax


def test_inner_slab_implies_the_fp64_slab_test():
    rng = np.random.default_rng(20261204)
    accepted = checked = 0
    for lo, hi, o, d, bmax in _cases(rng, 40000):
        checked += 1
        if inner_accepts(lo, hi, o, d, bmax):
            accepted += 1
            assert O.box_hit(lo, hi, o[None, :], d[None, :]), (lo, hi, o, d)
    # the inner test decides most boxes the rays are aimed at (it is not vacuous)
    assert checked > 39000 and accepted > 0.3 * checked


def test_inner_slab_rejects_a_missed_box_and_a_clamped_axis():
    lo = np.array([0.0, 0.0, 0.0], np.float32)
    hi = np.array([1.0, 1.0, 1.0], np.float32)
    o = np.array([0.5, 0.5, -2.0])
    # a ray through the middle: both accept it
    d1 = np.array([0.05, 0.1, 1.0]) / np.linalg.norm([0.05, 0.1, 1.0])
    assert inner_accepts(lo, hi, o, d1, 1.0)
    assert O.box_hit(lo, hi, o[None, :], d1[None, :])
    # the ray passes beside the box: neither test accepts it
    d = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    o2 = np.array([3.0, 0.5, -2.0])
    assert not inner_accepts(lo, hi, o2, d, 3.0)
    assert not O.box_hit(lo, hi, o2[None, :], d[None, :])
    # an axis with a zero direction component: 1/d is clamped to 1e30 there,
    # and the inner test accepts as long as the origin lies inside that slab
    # by more than the widening (about 6 * 2^-24 * (B + |o|))
    d0 = np.array([0.0, 0.0, 1.0])
    assert inner_accepts(lo, hi, o, d0, 1.0)
    assert O.box_hit(lo, hi, o[None, :], d0[None, :])
    # origin exactly on the slab's plane: the FP64 test takes 0 * inf as NaN
    # and ignores that axis (accepts); the inner test leaves it to FP64
    o_edge = np.array([0.0, 0.5, -2.0])
    assert not inner_accepts(lo, hi, o_edge, d0, 1.0)
    assert O.box_hit(lo, hi, o_edge[None, :], d0[None, :])
