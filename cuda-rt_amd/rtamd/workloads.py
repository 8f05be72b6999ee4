"""The benchmark workloads, as scenes on the host (inputs only; no oracle).

Each builder returns host meshes and the camera, light, material and
render options of one configuration. The meshes come from the library's
seeded C++ generators (rt_gen_soup, rt_gen_icosphere, rt_gen_jitter), so a
workload is the same on every machine.

- C1 ``c1_*``: a geodesic icosphere (radius 1, centre (0, 0, -4)) on a
  2-triangle ground quad, pinhole camera, 30-bit keys.
- C2 ``soup_mesh`` + ``c2_setup``: the random triangle soup (centroids
  uniform in the unit cube, lognormal edges of median 0.002 and sigma 0.5,
  random orientation) under a thin-lens camera (aperture 0.05, focus 1.5).
- C3 ``sphere10m_*``: a 9,996,980-triangle displaced icosphere (frequency
  707, 3-octave value noise of amplitude 0.05) with Gaussian vertex jitter
  per frame, refit only, 1024x1024 at 4 spp with diffuse bounces to depth 4.
- C4 ``batch_scenes``: 64 soups of log-uniform size in [1K, 2M] triangles,
  each with its own square camera of 256, 512 or 1024 pixels and 1..8 spp.
- C5 ``diff200k_*``: a 200,000-triangle displaced icosphere (frequency 100)
  with per-triangle albedo U[0.2, 0.9] and light intensity 10, 512x512 at
  16 spp with 2 diffuse bounces.
"""
import numpy as np

import rtamd

SEED_SOUP = 0x5EED0002
SEED_SPHERE10M = 0x5EED0003
SEED_BATCH = 0x5EED0004
SEED_DIFF = 0x5EED0006
SEED_DIFF_ALBEDO = 0x5EED0007


def c1_mesh(freq=32):
    """Icosphere (radius 1, centre (0,0,-4)) on a 2-triangle ground quad."""
    return rtamd.HostMesh.concat([rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -4.0)), rtamd.ground_quad()])


def c1_setup(width=512, height=512, spp=1):
    """Two Lambertian materials (kd 0.8), light at (2, 4, 0), 45-degree pinhole."""
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8)), rtamd.material(kd=(0.8, 0.8, 0.8))]
    light = rtamd.light(position=(2.0, 4.0, 0.0))
    cam = rtamd.camera(width, height, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    opts = rtamd.render_opts(spp)
    return mats, light, cam, opts


def soup_mesh(n, seed=SEED_SOUP):
    return rtamd.gen_soup(n, seed)


def c2_setup(width=1920, height=1080, spp=16):
    """Thin lens of aperture 0.05 (fstop 20) focused at 1.5, fov 40."""
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8))]
    light = rtamd.light(position=(1.0, 2.0, 0.5))
    cam = rtamd.camera(width, height, fov_deg=40.0, fstop=20.0, focal_point=1.5)
    opts = rtamd.render_opts(spp)
    return mats, light, cam, opts


def displaced_mesh(freq, seed=SEED_SPHERE10M):
    return rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -3.0), 0.05, seed)


def sphere10m_mesh(freq=707):
    return displaced_mesh(freq, SEED_SPHERE10M)


def sphere10m_frame(hm, k, sigma=1e-4):
    """Frame k's jittered positions (Gaussian, sigma 1e-4 per coordinate)."""
    return rtamd.gen_jitter(hm, sigma, 0x5EED1000 + k)


def sphere10m_setup(width=1024, height=1024, spp=4, depth=4):
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8))]
    light = rtamd.light(position=(2.0, 4.0, 0.0))
    cam = rtamd.camera(width, height, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    opts = rtamd.render_opts(spp, depth=depth, flags=rtamd.RT_RENDER_CLAMP)
    return mats, light, cam, opts


def batch_scenes(count=64, seed=SEED_BATCH, tri_range=(1_000, 2_000_000), resolutions=(256, 512, 1024),
                 spp_range=(1, 8), depth=0):
    """`count` random soups with log-uniform triangle counts, each with its
    own square camera and sample count (dicts of hm, mats, light, cam, opts)."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    for k in range(count):
        n = int(np.exp(rng.uniform(lo, hi)))
        res = int(rng.choice(resolutions))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        hm = rtamd.gen_soup(n, SEED_BATCH + 7919 * k)
        mats, light, cam, opts = c2_setup(res, res, spp)
        opts.depth = depth
        out.append(dict(hm=hm, mats=mats, light=light, cam=cam, opts=opts))
    return out


def diff200k_mesh(freq=100):
    return rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -3.0), 0.05, SEED_DIFF)


def diff200k_setup(width=512, height=512):
    light = rtamd.light(position=(2.0, 3.0, 0.0), color=(1.0, 1.0, 1.0), radius=0.5, c0=0.5, c1=0.2, c2=0.05)
    cam = rtamd.camera(width, height, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    return light, cam


def diff200k_albedo(n):
    """(true albedo U[0.2, 0.9], perturbed albedo: true + N(0, 0.05) clipped to [0, 1])."""
    rng = np.random.default_rng(SEED_DIFF_ALBEDO)
    true_alb = rng.uniform(0.2, 0.9, (n, 3))
    return true_alb, np.clip(true_alb + rng.normal(0, 0.05, true_alb.shape), 0.0, 1.0)
