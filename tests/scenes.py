"""Scene builders shared by the tests (small versions of the workloads)."""
import numpy as np

import rtamd


def c1_mesh(freq=32):
    """Icosphere (radius 1, centre (0,0,-4)) on a 2-triangle ground quad."""
    return rtamd.HostMesh.concat([rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -4.0)), rtamd.ground_quad()])


def soup_mesh(n, seed=0x5EED0002):
    return rtamd.gen_soup(n, seed)


def displaced_mesh(freq, seed=0x5EED0003):
    return rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -3.0), 0.05, seed)


def c1_setup(width=128, height=128, spp=1):
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8)), rtamd.material(kd=(0.8, 0.8, 0.8))]
    light = rtamd.light(position=(2.0, 4.0, 0.0))
    cam = rtamd.camera(width, height, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    opts = rtamd.render_opts(spp)
    return mats, light, cam, opts


def c2_setup(width, height, spp):
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8))]
    light = rtamd.light(position=(1.0, 2.0, 0.5))
    cam = rtamd.camera(width, height, fov_deg=40.0, fstop=20.0, focal_point=1.5)
    opts = rtamd.render_opts(spp)
    return mats, light, cam, opts


def random_rays(n, seed, origin=(0.0, 0.0, 0.0), spread=0.6):
    """Rays from near the origin towards -z with a random spread."""
    rng = np.random.default_rng(seed)
    o = np.tile(np.asarray(origin, np.float64), (n, 1)) + rng.normal(scale=0.01, size=(n, 3))
    d = np.stack([rng.uniform(-spread, spread, n), rng.uniform(-spread, spread, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    return o, d


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


def batch_scenes(count, seed, tri_range=(1_000, 2_000_000), resolutions=(256, 512, 1024), spp_range=(1, 8),
                 depth=0):
    """The batched workload: `count` random soups with log-uniform triangle
    counts, each with its own square camera and sample count."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    for k in range(count):
        n = int(np.exp(rng.uniform(lo, hi)))
        res = int(rng.choice(resolutions))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        hm = rtamd.gen_soup(n, 0x5EED0004 + 7919 * k)
        mats, light, cam, opts = c2_setup(res, res, spp)
        opts.depth = depth
        out.append(dict(hm=hm, mats=mats, light=light, cam=cam, opts=opts))
    return out
