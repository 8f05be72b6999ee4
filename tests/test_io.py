"""OBJ/MTL input, BMP output and the command-line front end, on the CPU.

The OBJ rules are checked on small hand-written files (each case names the
rule of the reference's tinyobj loader it pins) and on the logcabin scene
against the counts and bounds the reference printed for it
(results/logcabin.out, kept in tests/golden/logcabin_case.json). The BMP
writer's header must equal the reference image's header byte for byte.
"""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import rtamd

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "cuda-rt_amd", "rtamd_cli")


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.fixture(scope="module")
def logcabin(tmp_path_factory):
    d = tmp_path_factory.mktemp("logcabin")
    with gzip.open(os.path.join(GOLDEN, "logcabin.obj.gz"), "rb") as f:
        (d / "logcabin.obj").write_bytes(f.read())
    (d / "logcabin.mtl").write_bytes(open(os.path.join(GOLDEN, "logcabin.mtl"), "rb").read())
    return str(d / "logcabin.obj")


def test_logcabin_counts_bounds_materials(logcabin):
    case = json.load(open(os.path.join(GOLDEN, "logcabin_case.json")))
    hm, mats, info = rtamd.load_obj(logcabin)
    assert hm.n == case["counts"]["triangles"]
    assert info["raw_vertices"] == case["counts"]["vertices"]
    assert info["raw_normals"] == case["normals"]
    assert len(mats) == case["counts"]["materials"]
    assert info["shapes"] == len(case["shape_faces"])
    p = np.stack([hm.vx, hm.vy, hm.vz], 1)
    # the reference prints the float-rounded box bounds with %0.6f
    np.testing.assert_allclose(p.min(0), case["bounds_min"], atol=5e-7)
    np.testing.assert_allclose(p.max(0), case["bounds_max"], atol=5e-7)
    # '0.640000' accumulated digit by digit (0 + 6 * 10^-1 + 4 * 10^-2 + 0 ...)
    # is one ulp above the correctly rounded 0.64
    assert list(mats[0].kd) == [0.0 + 6 * 10.0 ** -1 + 4 * 10.0 ** -2] * 3 and mats[0].kd[0] != 0.64
    assert list(mats[1].kd) == [0.8] * 3 and list(mats[1].ks) == [0.8] * 3
    assert mats[0].shininess == pytest.approx(94.117647) and mats[1].shininess == 0.0
    # per-face material ids: 5348 faces of the first MTL material, 3990 of the second
    assert np.bincount(hm.mat_id).tolist() == [5348, 3990]
    # positions keep the file's doubles; the fp32 mesh is their rounding
    assert np.array_equal(info["positions64"].astype(np.float32), p)


def test_fan_triangulation_negative_indices_and_corner_forms(tmp_path):
    obj = write(tmp_path, "a.obj", """
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
vn 0 0 1
vn 0 1 0
vt 0 0
f 1 2 3 4
f -5/1/1 -4//2 -1/1
""")
    hm, mats, info = rtamd.load_obj(obj)
    assert hm.n == 3
    tri = hm.idx.reshape(-1, 3)
    pos = np.stack([hm.vx, hm.vy, hm.vz], 1)
    # fan (0,1,2), (0,2,3) of the quad; then the triangle of vertices 1, 2, 5
    got = [[tuple(pos[k]) for k in t] for t in tri]
    assert got[0] == [(0, 0, 0), (1, 0, 0), (1, 1, 0)]
    assert got[1] == [(0, 0, 0), (1, 1, 0), (0, 1, 0)]
    assert got[2] == [(0, 0, 0), (1, 0, 0), (0, 0, 1)]
    nrm = np.stack([hm.nx, hm.ny, hm.nz], 1)
    # corner normals follow their own index; a corner without one gets 0
    assert [tuple(nrm[k]) for k in tri[2]] == [(0, 0, 1), (0, 1, 0), (0, 0, 0)]
    # one vertex per distinct (position, normal) pair: the quad's 4 corners
    # have no normal, the triangle reuses positions 1 and 2 with normals
    assert hm.nv == 7
    assert hm.mat_id.tolist() == [-1, -1, -1]


def test_usemtl_grouping_and_dropped_group(tmp_path):
    write(tmp_path, "m.mtl", """newmtl red
Kd 1 0 0
newmtl blue
Kd 0 0 1
Ns 10
""")
    obj = write(tmp_path, "b.obj", """mtllib m.mtl
v 0 0 0
v 1 0 0
v 0 1 0
usemtl red
f 1 2 3
usemtl blue
g second
f 1 3 2
usemtl nosuchmaterial
f 2 3 1
""")
    hm, mats, info = rtamd.load_obj(obj)
    # 'usemtl blue' flushes the red face into the open shape; 'g' then finds
    # no pending face and drops that shape (tinyobj's rule), so only the
    # faces after 'g' survive: blue, then the unknown material (-1).
    assert hm.n == 2
    assert hm.mat_id.tolist() == [1, -1]
    assert [list(m.kd) for m in mats] == [[1, 0, 0], [0, 0, 1]]
    assert mats[0].shininess == 1.0 and mats[1].shininess == 10.0


def test_number_syntax_and_missing_mtl(tmp_path):
    obj = write(tmp_path, "c.obj", """mtllib missing.mtl
v .5 -0.25 1e2
v +2.5E-1 3 -4.
v 1 1 1x
f 1 2 3
""")
    hm, mats, info = rtamd.load_obj(obj)
    p = info["positions64"]
    # no leading digit -> the default 0; exponents; a trailing letter ends the number
    assert p[0].tolist() == [0.0, -0.25, 100.0]
    assert p[1].tolist() == [0.25, 3.0, -4.0]
    assert p[2].tolist() == [1.0, 1.0, 1.0]
    # a missing MTL file still yields one default material (shininess 1)
    assert len(mats) == 1 and list(mats[0].kd) == [0, 0, 0] and mats[0].shininess == 1.0


def test_bmp_header_and_pixels(tmp_path):
    ref = np.load(os.path.join(GOLDEN, "logcabin_reference_image.npz"))
    out = str(tmp_path / "x.bmp")
    rtamd.write_bmp(out, 1024, 1024, np.zeros(1024 * 1024 * 3, np.float32))
    data = open(out, "rb").read()
    assert data[:54] == ref["header"].tobytes()
    assert len(data) == 54 + 1024 * 1024 * 3
    # 3 x 2 image: rows padded to 12 bytes, BGR, clamp and truncating quantisation
    rgb = np.array([[0.0, 0.5, 1.0], [2.0, -1.0, 0.999], [0.2, 0.4, 0.6],
                    [1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [0.1, 0.2, 0.3]], np.float32)
    out = str(tmp_path / "y.bmp")
    rtamd.write_bmp(out, 3, 2, rgb)
    d = open(out, "rb").read()
    assert int.from_bytes(d[2:6], "little") == 56 + 24
    px = np.frombuffer(d[54:], np.uint8).reshape(2, 12)
    q = (255.0 * np.clip(rgb.astype(np.float64), 0, 1)).astype(np.uint8).reshape(2, 3, 3)[:, :, ::-1]
    assert np.array_equal(px[:, :9].reshape(2, 3, 3), q)
    assert (px[:, 9:] == 0).all()


def test_cli_help_and_bad_flag():
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--focalpoint" in r.stdout
    r = subprocess.run([CLI, "--no-such-flag"], capture_output=True, text=True)
    assert r.returncode == 1
