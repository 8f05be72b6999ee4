"""Scene builders for the tests: the workloads of rtamd.workloads at test sizes,
plus test-only rays and adversarial meshes."""
import numpy as np

import rtamd
from rtamd import workloads as _w
from rtamd.workloads import batch_scenes, c1_mesh, c2_setup, displaced_mesh, soup_mesh  # noqa: F401


def c1_setup(width=128, height=128, spp=1):
    """C1's materials, light and pinhole camera at a test-sized image."""
    return _w.c1_setup(width, height, spp)


def random_rays(n, seed, origin=(0.0, 0.0, 0.0), spread=0.6):
    """Rays from near the origin towards -z with a random spread."""
    rng = np.random.default_rng(seed)
    o = np.tile(np.asarray(origin, np.float64), (n, 1)) + rng.normal(scale=0.01, size=(n, 3))
    d = np.stack([rng.uniform(-spread, spread, n), rng.uniform(-spread, spread, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    return o, d


def deep_chain_mesh(hittable=(), bits=63):
    """Flat triangles whose Morton keys are C + (2^i - 1), i = 0..54 (C: the
    top bit of each axis), plus degenerate triangles at the scene corners
    (keys 0 and all ones). Consecutive keys then first differ at ever higher
    bits, so the radix tree is a chain about 57 levels deep. Every box
    contains the line x = y = 0.499; a triangle covers x + y >= cx + cy >= 1
    (so the line misses it) unless its index is in `hittable`, where it
    covers x + y <= cx + cy. Returns (mesh, index of each chain triangle)."""
    assert bits == 63
    h = 0.249
    tris = [np.zeros((3, 3)), np.ones((3, 3))]
    for i in range(55):
        # bit 3k of a key is z's bit k, 3k+1 y's, 3k+2 x's (kernels/bvh.cu:31)
        j = [i // 3, i // 3 + (1 if i % 3 >= 2 else 0), i // 3 + (1 if i % 3 >= 1 else 0)]  # x, y, z
        c = [((1 << 20) + (1 << jj) - 1 + 0.5) / float(1 << 21) for jj in j]
        cx, cy, cz = c
        if i in hittable:
            v = [(cx - h, cy - h, cz), (cx + h, cy - h, cz), (cx - h, cy + h, cz)]
        else:
            v = [(cx + h, cy + h, cz), (cx - h, cy + h, cz), (cx + h, cy - h, cz)]
        tris.append(np.array(v))
    t = np.array(tris, np.float64)
    n = len(t)
    p = t.reshape(-1, 3).astype(np.float32)
    nrm = np.tile(np.array([0.0, 0.0, 1.0], np.float32), (3 * n, 1))
    idx = np.arange(3 * n, dtype=np.int32)
    hm = rtamd.HostMesh(p[:, 0], p[:, 1], p[:, 2], nrm[:, 0], nrm[:, 1], nrm[:, 2], idx)
    return hm, np.arange(2, n)


def duplicate_key_mesh(n, seed=77):
    """Most triangles packed into one tiny cluster, a few spread over the unit
    cube: the cluster's centroids share Morton cells, so keys repeat and the
    tree falls back to index bits (kernels/bvh.cu:87)."""
    hm = rtamd.gen_soup(n, seed, edge_median=1e-4)
    k = n - 16
    c = np.float32(0.3)
    for a in (hm.vx, hm.vy, hm.vz):
        a[: 3 * k] = c + (a[: 3 * k] - a[: 3 * k].mean()) * np.float32(1e-3)
        a[3: 3 * 1000] = np.tile(a[0:3], 999)  # 1000 identical triangles: equal keys at any width
    return hm
