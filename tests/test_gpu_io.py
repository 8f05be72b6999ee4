"""The OBJ -> build -> render -> BMP path on the GPU: the command-line
renderer on the logcabin scene with the flags of runLogcabin.sh:3, against
the oracle and against the reference's own output image
(results/logcabin.bmp, kept in tests/golden), and the supersampling resolve.
"""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import rtamd
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "cuda-rt_amd", "rtamd_cli")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return rtamd.Context(0)


@pytest.fixture(scope="module")
def logcabin(tmp_path_factory):
    d = tmp_path_factory.mktemp("logcabin")
    with gzip.open(os.path.join(GOLDEN, "logcabin.obj.gz"), "rb") as f:
        (d / "logcabin.obj").write_bytes(f.read())
    (d / "logcabin.mtl").write_bytes(open(os.path.join(GOLDEN, "logcabin.mtl"), "rb").read())
    return d


def _read_bmp(path):
    raw = open(path, "rb").read()
    return raw[:54], np.frombuffer(raw[54:], np.uint8).reshape(1024, 1024, 3)


def test_cli_logcabin_matches_oracle_and_reference_image(ctx, logcabin):
    _say("CLI logcabin 1024x1024 x20spp (runLogcabin.sh flags)")
    case = json.load(open(os.path.join(GOLDEN, "logcabin_case.json")))
    args = list(case["args"])
    args[args.index("-i") + 1] = str(logcabin / "logcabin.obj")
    out = str(logcabin / "ours.bmp")
    args[args.index("-o") + 1] = out
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "num of triangles : 9338" in r.stdout
    assert "found min(-6.816752, -2.977815 , -24.966003)found max(9.500026, 3.785227 , -8.857635)" in r.stdout
    header, ours = _read_bmp(out)
    ref = np.load(os.path.join(GOLDEN, "logcabin_reference_image.npz"))
    assert header == ref["header"].tobytes()

    # the same frame through the oracle: FP64 restatement, quantised the same way
    hm, mats, info = rtamd.load_obj(str(logcabin / "logcabin.obj"))
    f32 = lambda x: float(np.float32(x))
    light = rtamd.light(position=(f32(-4.0), f32(2.0), f32(-4.0)), color=(1.0, 1.0, 1.0), radius=0.5,
                        c0=f32(0.0001), c1=f32(0.0001), c2=f32(0.01))
    cam = rtamd.camera(1024, 1024, fov_deg=35.0, fstop=16.0, focal_point=15.0)
    opts = rtamd.render_opts(20, depth=0, seed=0, flags=rtamd.RT_RENDER_CLAMP)
    rgb, _ = O.render(hm, O.build(hm, 30), mats, light, cam, opts, nthreads=16)
    q = (255.0 * np.clip(rgb.astype(np.float64), 0, 1)).astype(np.uint8).reshape(1024, 1024, 3)[:, :, ::-1]
    d = np.abs(ours.astype(int) - q.astype(int))
    assert d.max() <= 1, d.max()
    assert (d == 0).all(2).mean() >= 0.999

    # against the reference's own image. Its random stream is not reproduced
    # (seed 0 matches the draw-sensitive pixels no better than any other
    # seed, tests/test_logcabin_stats.py), so per pixel the image is checked
    # as one more draw of our renderer: its deviation from the mean of an
    # 8-seed ensemble of GPU renders equals that of held-out seeds.
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, 30)
    mt = rtamd.materials_tensor(mats)

    def gpu_q(seed):
        o = rtamd.render_opts(20, depth=0, seed=seed, flags=rtamd.RT_RENDER_CLAMP | rtamd.RT_RENDER_WAVEFRONT)
        rgb = rtamd.render(ctx, bvh, dm, mt, len(mats), light, cam, o)
        torch.cuda.synchronize()
        return (255.0 * np.clip(rgb.cpu().numpy().astype(np.float64), 0, 1)).astype(np.uint8).reshape(1024, 1024, 3)

    ens = np.stack([gpu_q(s) for s in range(1, 9)]).astype(float)
    live = (ens.max(0) != ens.min(0)).any(2)
    mu = ens.mean(0)[live]
    want = ref["bgr"][:, :, ::-1].astype(float)[live]
    held = [gpu_q(s).astype(float)[live] for s in (101, 102)]
    mad = np.abs(want - mu).mean() / np.mean([np.abs(h - mu).mean() for h in held])
    mse = ((want - mu) ** 2).mean() / np.mean([((h - mu) ** 2).mean() for h in held])
    _say(f"reference image vs GPU ensemble: live {live.mean():.4f} MAD ratio {mad:.4f} MSE ratio {mse:.4f}")
    assert live.mean() > 0.05
    assert 0.95 <= mad <= 1.05 and 0.92 <= mse <= 1.08, (mad, mse)
    # where the draws do not move the pixel the images agree (measured on the
    # oracle: 99.2 % exact, 99.99 % within 1 LSB)
    dr = np.abs(ours.astype(int) - ref["bgr"].astype(int)).max(2)
    assert (dr[~live] == 0).mean() >= 0.985 and (dr[~live] <= 1).mean() >= 0.999
    assert abs(ours.astype(float).mean() - ref["bgr"].astype(float).mean()) < 0.05
    bvh.close()


@pytest.mark.parametrize("ss", [1, 2, 3])
def test_supersample_resolve(ctx, ss):
    _say(f"supersampling resolve ss={ss}")
    W, H = 40, 24
    g = torch.Generator().manual_seed(ss)
    big = (torch.rand(W * ss * H * ss * 3, generator=g, dtype=torch.float32) * 1.6 - 0.3).cuda()
    b = big.cpu().numpy().astype(np.float64).reshape(H * ss, W * ss, 3)
    blocks = np.clip(b, 0, 1).reshape(H, ss, W, ss, 3).sum((1, 3))
    for flags, div in ((0, ss * ss), (rtamd.RT_RESOLVE_REFERENCE_DIVISOR, ss)):
        got = rtamd.resolve_supersample(ctx, big, W, H, ss, flags).cpu().numpy().reshape(H, W, 3)
        want = blocks if ss == 1 else blocks / div
        assert np.array_equal(got, want.astype(np.float32))
