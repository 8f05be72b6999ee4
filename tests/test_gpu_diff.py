"""Differentiable render: GPU forward against the FP64 CPU oracle, GPU adjoint
gradients against central finite differences of the GPU forward and of the
oracle's forward.

Albedo and light intensity enter the image polynomially, so central FD is
exact up to rounding and the bar is 1e-6 relative. Positions move hit
points, normals and light directions; the adjoint holds visibility fixed,
so the FD loss is restricted to pixels whose path (hit leaf, occluder and
shading branch of every bounce) is the same at every perturbed position and
whose difference is linear in the step. The bar there is 1e-3 relative.
"""
import numpy as np
import pytest

import rtamd
from rtamd import diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


def _scene(freq=10, res=48, depth=2):
    hm = rtamd.HostMesh.concat([rtamd.gen_icosphere(freq, 1.0, (0.0, 0.0, -3.0), 0.05, 0x5EED0005),
                                rtamd.ground_quad(y=-1.2)])
    light = rtamd.light(position=(2.0, 4.0, 0.0), color=(1.0, 0.9, 0.8), radius=0.5, c0=0.5, c1=0.2, c2=0.05)
    cam = rtamd.camera(res, res, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    return hm, light, cam


def _renderer(ctx, hm, light, cam, spp=4, depth=2, seed=3):
    dm = rtamd.DeviceMesh(hm)
    return diff.DiffRenderer(ctx, dm, light, cam, spp, depth, seed=seed)


def _albedo(n, seed=11):
    rng = np.random.default_rng(seed)
    return torch.tensor(rng.uniform(0.2, 0.9, (n, 3)), dtype=torch.float64, device="cuda")


def _oracle_render(hm, albedo, s, light, cam, spp, depth, seed, positions=None, branches=False):
    """The same scene through the CPU oracle: one material per triangle."""
    n = hm.n
    vx, vy, vz = hm.vx, hm.vy, hm.vz
    if positions is not None:
        vx, vy, vz = (np.ascontiguousarray(positions[:, k], np.float32) for k in range(3))
    m = rtamd.HostMesh(vx, vy, vz, hm.nx, hm.ny, hm.nz, hm.idx, np.arange(n, dtype=np.int32))
    a = albedo.cpu().numpy()
    mats = [rtamd.material(kd=tuple(a[t])) for t in range(n)]
    L = rtamd.light(position=tuple(light.position), color=tuple(s * c for c in light.color), radius=light.radius,
                    c0=light.c0, c1=light.c1, c2=light.c2)
    opts = rtamd.render_opts(spp, depth=depth, seed=seed, flags=0)
    b = O.build(m, 30)
    out = O.render(m, b, mats, L, cam, opts, fp64=True, records=True, branches=branches)
    return out[2:]


def _leaf_to_tri(rec, ids):
    out = rec.copy()
    ids = np.asarray(ids, np.int64)
    for c in (0, 1):
        v = out[..., c]
        hit = v >= 0
        v[hit] = ids[v[hit]]
    return out


def _positions(hm):
    return np.stack([hm.vx, hm.vy, hm.vz], 1).astype(np.float32)


def _pixel_mask(*recs):
    """True for pixels whose records agree across all given record arrays."""
    base = recs[0].reshape(recs[0].shape[0], -1)
    same = np.ones(base.shape[0], bool)
    for r in recs[1:]:
        same &= (r.reshape(r.shape[0], -1) == base).all(1)
    return same


def _image_mask(cam, pix_mask):
    """Map a launch-pixel mask (py*width+px) to the row-flipped image layout."""
    W, H = cam.width, cam.height
    m = pix_mask.reshape(H, W)[::-1]
    return np.repeat(m.reshape(-1), 3)


def test_diff_forward_matches_oracle(ctx):
    _say("diff forward parity: displaced sphere freq 10 + ground, 48x48 x4spp depth 2")
    hm, light, cam = _scene()
    R = _renderer(ctx, hm, light, cam)
    albedo = _albedo(hm.n)
    s = 10.0
    img = R.forward(albedo, s)
    torch.cuda.synchronize()
    got = img.cpu().numpy().reshape(-1)
    ref, rec_ref = _oracle_render(hm, albedo, s, light, cam, 4, 2, 3)
    d = np.abs(got - ref)
    assert d.max() <= 1e-4, f"max image diff {d.max()}"
    assert got.max() > 0.1
    ids = R.bvh.export()["ids"]
    rec = R.records.cpu().numpy().reshape(rec_ref.shape)
    assert np.array_equal(_leaf_to_tri(rec, ids), rec_ref), "path records differ"
    # the scene exercises misses, shadowed and unshadowed hits and bounces
    assert (rec_ref[..., 0, 0] == -1).any() and (rec_ref[..., 1, 0] >= 0).any()
    assert (rec_ref[..., 1] >= 0).any() and (rec_ref[..., 0, 1] == -1).any()


def _loss(img, w):
    return float((img * w).sum())


def test_diff_albedo_light_gradients_fd(ctx):
    _say("diff albedo/light gradients vs central FD of the GPU forward")
    hm, light, cam = _scene()
    R = _renderer(ctx, hm, light, cam)
    albedo = _albedo(hm.n)
    s = 10.0
    rng = np.random.default_rng(5)
    w = torch.tensor(rng.normal(size=R.shape), dtype=torch.float64, device="cuda")
    R.forward(albedo, s)
    g_alb, g_light, _ = R.backward(albedo, s, w)
    torch.cuda.synchronize()
    # light intensity (the image is linear in it)
    h = 1e-3
    fd = (_loss(R.forward(albedo, s + h), w) - _loss(R.forward(albedo, s - h), w)) / (2 * h)
    assert abs(float(g_light) - fd) <= 1e-6 * abs(fd) + 1e-9, (float(g_light), fd)
    # random direction over every albedo
    v = torch.tensor(rng.normal(size=(hm.n, 3)), dtype=torch.float64, device="cuda")
    h = 1e-4
    fd = (_loss(R.forward(albedo + h * v, s), w) - _loss(R.forward(albedo - h * v, s), w)) / (2 * h)
    pred = float((g_alb * v).sum())
    assert abs(pred - fd) <= 1e-6 * abs(fd) + 1e-9, (pred, fd)
    # single entries: the largest gradients (directly lit, bounce and occluder terms)
    ga = g_alb.cpu().numpy()
    for t, c in zip(*np.unravel_index(np.argsort(-np.abs(ga).ravel())[:6], ga.shape)):
        e = torch.zeros_like(albedo)
        e[t, c] = 1.0
        fd = (_loss(R.forward(albedo + h * e, s), w) - _loss(R.forward(albedo - h * e, s), w)) / (2 * h)
        assert abs(ga[t, c] - fd) <= 1e-6 * abs(fd) + 1e-9, (t, c, ga[t, c], fd)
    # a triangle never seen has no gradient
    assert np.count_nonzero(np.abs(ga).sum(1) == 0) > 0


def _oracle_paths(hm, albedo, s, light, cam, pos, spp=4, depth=2, seed=3):
    """Oracle image and, per pixel, its path record plus shading branch bits."""
    img, rec, br = _oracle_render(hm, albedo, s, light, cam, spp, depth, seed, positions=pos, branches=True)
    npix = cam.width * cam.height
    return img, np.concatenate([rec.reshape(npix, -1), br.reshape(npix, -1)], 1)


def _fd_positions(images, paths, pts, cam, seed):
    """Finite-difference check data along one direction of the positions.

    pts are the positions at steps +h, -h, +h/2, -h/2 and 0 (float-rounded),
    images the flat images there and paths the per-pixel path signatures
    (hit leaves, occluders and shading branches: N.L > 0, falloff clamp,
    normal flip, bounce frame axis). Kept are the values of pixels whose
    signature is the same at all five positions (so the forward is smooth
    between them) and whose difference is linear in the step
    (|D(h) - 2 D(h/2)| within 2e-3 relative: no steep curvature). The
    weights are +-U(0.5, 1.5) with the sign of D(h/2), so the kept
    differences add up instead of cancelling. Returns the weights,
    Richardson's (4 S(h/2) - S(h)) / 3 of the weighted sums S, and the two
    exact position differences.
    """
    rng = np.random.default_rng(seed)
    keep_pix = _pixel_mask(*paths)
    d1 = images[0] - images[1]
    d2 = images[2] - images[3]
    lin = (np.abs(d1 - 2 * d2) <= 1e-7 + 2e-3 * np.abs(d1)).reshape(-1, 3).all(1)
    keep = _image_mask(cam, keep_pix) & np.repeat(lin, 3)
    w = np.sign(d2) * keep * rng.uniform(0.5, 1.5, size=keep.shape)
    fd = (4 * float((w * d2).sum()) - float((w * d1).sum())) / 3
    dp1 = pts[0].astype(np.float64) - pts[1].astype(np.float64)
    dp2 = pts[2].astype(np.float64) - pts[3].astype(np.float64)
    return w, keep, fd, dp1, dp2


def _pos_setup(ctx, seed, h=5e-4):
    hm, light, cam = _scene()
    R = _renderer(ctx, hm, light, cam)
    albedo = _albedo(hm.n)
    rng = np.random.default_rng(seed)
    p0 = _positions(hm)
    v = rng.normal(size=p0.shape)
    pts = [(p0 + f * h * v).astype(np.float32) for f in (1.0, -1.0, 0.5, -0.5, 0.0)]
    return hm, light, cam, R, albedo, pts


def _adjoint_pred(R, albedo, s, p0, w, dp1, dp2):
    R.set_positions(torch.from_numpy(p0).cuda())
    R.forward(albedo, s)
    _, _, g_pos = R.backward(albedo, s, torch.from_numpy(w.reshape(R.shape)).cuda())
    g = g_pos.cpu().numpy()
    return (4 * float((g * dp2).sum()) - float((g * dp1).sum())) / 3


@pytest.mark.parametrize("seed", [21, 22])
def test_diff_position_gradients_fd(ctx, seed):
    _say(f"diff position gradients vs central FD of the GPU forward, seed {seed}")
    hm, light, cam, R, albedo, pts = _pos_setup(ctx, seed)
    s = 10.0
    npix = cam.width * cam.height
    images, paths = [], []
    for pos in pts:
        R.set_positions(torch.from_numpy(pos).cuda())
        images.append(R.forward(albedo, s).cpu().numpy().reshape(-1))
        paths.append(np.concatenate([R.records.cpu().numpy().reshape(npix, -1),
                                     _oracle_paths(hm, albedo, s, light, cam, pos)[1]], 1))
    w, keep, fd, dp1, dp2 = _fd_positions(images, paths, pts, cam, seed)
    assert keep.mean() > 0.7, keep.mean()
    pred = _adjoint_pred(R, albedo, s, pts[4], w, dp1, dp2)
    assert abs(pred - fd) <= 1e-3 * abs(fd), (pred, fd)


def test_diff_position_gradients_vs_oracle_fd(ctx):
    _say("diff position gradients vs central FD of the FP64 CPU oracle")
    hm, light, cam, R, albedo, pts = _pos_setup(ctx, 31)
    s = 10.0
    out = [_oracle_paths(hm, albedo, s, light, cam, pos) for pos in pts]
    w, keep, fd, dp1, dp2 = _fd_positions([o[0] for o in out], [o[1] for o in out], pts, cam, 31)
    assert keep.mean() > 0.7, keep.mean()
    pred = _adjoint_pred(R, albedo, s, pts[4], w, dp1, dp2)
    assert abs(pred - fd) <= 1e-3 * abs(fd), (pred, fd)


def test_diff_autograd_function(ctx):
    _say("diff autograd function: loss.backward matches rt_diff_backward")
    hm, light, cam = _scene(freq=6, res=32)
    R = _renderer(ctx, hm, light, cam, spp=2, depth=1)
    albedo = _albedo(hm.n).requires_grad_(True)
    s = torch.tensor(8.0, dtype=torch.float64, device="cuda", requires_grad=True)
    pos = torch.from_numpy(_positions(hm).astype(np.float64)).cuda().requires_grad_(True)
    target = torch.zeros(R.shape, dtype=torch.float64, device="cuda")
    img = diff.render(R, albedo, s, pos)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    g_alb, g_light, g_pos = R.backward(albedo, 8.0, 2 * (img.detach() - target))
    torch.testing.assert_close(albedo.grad, g_alb)
    torch.testing.assert_close(s.grad, g_light)
    torch.testing.assert_close(pos.grad, g_pos)
    assert float(albedo.grad.abs().sum()) > 0 and float(pos.grad.abs().sum()) > 0


def test_diff_rejects_bad_arguments(ctx):
    hm, light, cam = _scene(freq=4, res=16)
    with pytest.raises(ValueError):
        _renderer(ctx, hm, light, cam, spp=1, depth=diff.RT_DIFF_MAX_DEPTH + 1)
    R = _renderer(ctx, hm, light, cam, spp=1, depth=1)
    with pytest.raises(rtamd.RtError):
        R.forward(_albedo(hm.n), -1.0)
    with pytest.raises(ValueError):
        R.forward(_albedo(hm.n + 1), 1.0)
