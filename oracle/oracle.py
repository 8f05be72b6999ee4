"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: a double-precision restatement of the reference
algorithm (mbartling/cuda-rt) used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker and CPU baseline. The product
(cuda-rt_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

vp = ctypes.c_void_p
c_f = ctypes.POINTER(ctypes.c_float)
c_d = ctypes.POINTER(ctypes.c_double)
c_i32 = ctypes.POINTER(ctypes.c_int32)
c_u32 = ctypes.POINTER(ctypes.c_uint32)
c_u64 = ctypes.POINTER(ctypes.c_uint64)
c_i64 = ctypes.POINTER(ctypes.c_int64)

RANDVEC_RTL, CLAMP = 1, 2
# planted mistakes (tests only): light jitter (rand - 0.5) * radius; aperture 1.25 / fstop; lens radius without sqrt
PLANT_LIGHT_OFFSET, PLANT_APERTURE, PLANT_LENS_RADIUS = 1 << 24, 1 << 25, 1 << 26


class Mesh(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nv", ctypes.c_int32), ("vx", c_f), ("vy", c_f), ("vz", c_f),
                ("nx", c_f), ("ny", c_f), ("nz", c_f), ("idx", c_i32), ("mat_id", c_i32), ("pos64", c_d)]


class Bvh(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("parent", c_i32), ("left", c_i32), ("right", c_i32),
                ("lo", c_f), ("hi", c_f)]


class Material(ctypes.Structure):
    _fields_ = [("ka", ctypes.c_double * 3), ("kd", ctypes.c_double * 3), ("ks", ctypes.c_double * 3),
                ("kt", ctypes.c_double * 3), ("ke", ctypes.c_double * 3), ("shininess", ctypes.c_double)]


class Light(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("color", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("c0", ctypes.c_double), ("c1", ctypes.c_double), ("c2", ctypes.c_double)]


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fov_deg", ctypes.c_double),
                ("fstop", ctypes.c_double), ("focal_point", ctypes.c_double), ("eye", ctypes.c_double * 3)]


class RenderOpts(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_int32), ("depth", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("flags", ctypes.c_int32)]


_lib = None


def compile_lib():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            compile_lib()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_wang_hash.restype = ctypes.c_uint32
        L.orc_wang_hash.argtypes = [ctypes.c_uint32]
        L.orc_xorshift.restype = ctypes.c_uint32
        L.orc_xorshift.argtypes = [c_u32]
        L.orc_rand_double.restype = ctypes.c_double
        L.orc_rand_double.argtypes = [c_u32]
        L.orc_seed_index.restype = ctypes.c_uint32
        L.orc_seed_index.argtypes = [ctypes.c_int32] * 3
        L.orc_expand_bits10.restype = ctypes.c_uint32
        L.orc_expand_bits10.argtypes = [ctypes.c_uint32]
        L.orc_sincos.argtypes = [ctypes.c_double, c_d, c_d]
        L.orc_cosine_dir.argtypes = [c_d, ctypes.c_double, ctypes.c_double, c_d]
        L.orc_box_hit.restype = ctypes.c_int32
        L.orc_intersect_triangle.restype = ctypes.c_int32
        # entry points with 64-bit sizes need explicit argtypes
        L.orc_trace.argtypes = [vp, vp, ctypes.c_int64, c_d, c_d, ctypes.c_int32, c_i32, c_d, c_d, c_d, c_d, c_i64,
                                ctypes.c_int32]
        L.orc_trace_brute.argtypes = [vp, c_i32, ctypes.c_int64, c_d, c_d, c_i32, c_d]
        L.orc_render.argtypes = [vp, vp, vp, ctypes.c_int32, vp, vp, vp, c_f, c_i64, ctypes.c_int64, c_i64,
                                 ctypes.c_int32]
        L.orc_render64.argtypes = [vp, vp, vp, ctypes.c_int32, vp, vp, vp, c_f, c_d, c_i32, c_i32, c_i64,
                                   ctypes.c_int64, c_i64, ctypes.c_int32]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ct)


class OMesh:
    """Keeps numpy arrays alive behind an orc_mesh."""

    def __init__(self, hm):
        self.hm = hm
        self.mat = hm.mat_id
        p64 = getattr(hm, "pos64", None)
        self.p64 = None if p64 is None else np.ascontiguousarray(p64, np.float64)
        self.s = Mesh(hm.n, hm.nv, _p(hm.vx, c_f), _p(hm.vy, c_f), _p(hm.vz, c_f), _p(hm.nx, c_f), _p(hm.ny, c_f),
                      _p(hm.nz, c_f), _p(hm.idx, c_i32), _p(hm.mat_id, c_i32) if hm.mat_id is not None else None,
                      _p(self.p64, c_d) if self.p64 is not None else None)


def tri_aabb(hm):
    om = OMesh(hm)
    lo = np.empty((hm.n, 3), np.float32)
    hi = np.empty((hm.n, 3), np.float32)
    lib().orc_tri_aabb(ctypes.byref(om.s), _p(lo, c_f), _p(hi, c_f))
    return lo, hi


def scene_bounds(lo, hi):
    mn = np.empty(3)
    mx = np.empty(3)
    lib().orc_scene_bounds(lo.shape[0], _p(np.ascontiguousarray(lo), c_f), _p(np.ascontiguousarray(hi), c_f),
                           _p(mn, c_d), _p(mx, c_d))
    return mn, mx


def morton(lo, hi, mn, mx, bits):
    n = lo.shape[0]
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    mn = np.ascontiguousarray(mn, np.float64)
    mx = np.ascontiguousarray(mx, np.float64)
    if bits == 30:
        c = np.empty(n, np.uint32)
        lib().orc_morton30(n, _p(lo, c_f), _p(hi, c_f), _p(mn, c_d), _p(mx, c_d), _p(c, c_u32))
    else:
        c = np.empty(n, np.uint64)
        lib().orc_morton63(n, _p(lo, c_f), _p(hi, c_f), _p(mn, c_d), _p(mx, c_d), _p(c, c_u64))
    return c


def sort(codes):
    c = codes.copy()
    ids = np.empty(c.size, np.uint32)
    if c.dtype == np.uint32:
        lib().orc_sort30(c.size, _p(c, c_u32), _p(ids, c_u32))
    else:
        lib().orc_sort63(c.size, _p(c, c_u64), _p(ids, c_u32))
    return c, ids


def hierarchy(codes, ids, topdown=False):
    n = codes.size
    parent = np.empty(2 * n - 1, np.int32)
    left = np.empty(2 * n - 1, np.int32)
    right = np.empty(2 * n - 1, np.int32)
    ids = np.ascontiguousarray(ids, np.uint32)
    kind = ("topdown" if topdown else "hierarchy") + ("30" if codes.dtype == np.uint32 else "63")
    fn = getattr(lib(), "orc_" + kind)
    fn(n, _p(codes, c_u32 if codes.dtype == np.uint32 else c_u64), _p(ids, c_u32), _p(parent, c_i32),
       _p(left, c_i32), _p(right, c_i32))
    return parent, left, right


def refit(parent, left, right, tri_lo, tri_hi, perm=None, brute=False):
    n = (parent.size + 1) // 2
    lo = np.empty((2 * n - 1, 3), np.float32)
    hi = np.empty((2 * n - 1, 3), np.float32)
    tri_lo = np.ascontiguousarray(tri_lo, np.float32)
    tri_hi = np.ascontiguousarray(tri_hi, np.float32)
    if brute:
        lib().orc_refit_brute(n, _p(left, c_i32), _p(right, c_i32), _p(tri_lo, c_f), _p(tri_hi, c_f),
                              _p(lo, c_f), _p(hi, c_f))
    else:
        pp = None if perm is None else _p(np.ascontiguousarray(perm, np.int32), c_i32)
        lib().orc_refit(n, _p(parent, c_i32), _p(left, c_i32), _p(right, c_i32), _p(tri_lo, c_f), _p(tri_hi, c_f),
                        pp, _p(lo, c_f), _p(hi, c_f))
    return lo, hi


def build(hm, bits=30, nthreads=0):
    """Full oracle build; returns a dict in the reference layout."""
    n = hm.n
    om = OMesh(hm)
    codes = np.empty(n, np.uint64)
    ids = np.empty(n, np.uint32)
    parent = np.empty(2 * n - 1, np.int32)
    left = np.empty(2 * n - 1, np.int32)
    right = np.empty(2 * n - 1, np.int32)
    lo = np.empty((2 * n - 1, 3), np.float32)
    hi = np.empty((2 * n - 1, 3), np.float32)
    mn = np.empty(3)
    mx = np.empty(3)
    lib().orc_build(ctypes.byref(om.s), bits, _p(codes, c_u64), _p(ids, c_u32), _p(parent, c_i32), _p(left, c_i32),
                    _p(right, c_i32), _p(lo, c_f), _p(hi, c_f), _p(mn, c_d), _p(mx, c_d), nthreads)
    if bits == 30:
        codes = codes.astype(np.uint32)
    return dict(codes=codes, ids=ids, parent=parent, left=left, right=right, lo=lo, hi=hi, mn=mn, mx=mx)


class OBvh:
    def __init__(self, b):
        self.b = b
        self.s = Bvh(b["parent"].size // 2 + 1, _p(b["parent"], c_i32), _p(b["left"], c_i32), _p(b["right"], c_i32),
                     _p(b["lo"], c_f), _p(b["hi"], c_f))


def trace(hm, b, orig, dirs, any_hit=False, nthreads=0):
    om = OMesh(hm)
    ob = OBvh(b)
    n = orig.shape[0]
    orig = np.ascontiguousarray(orig, np.float64)
    dirs = np.ascontiguousarray(dirs, np.float64)
    tri = np.empty(n, np.int32)
    t = np.empty(n)
    a = np.empty(n)
    be = np.empty(n)
    nrm = np.empty((n, 3))
    visits = np.empty((n, 2), np.int64)
    lib().orc_trace(ctypes.byref(om.s), ctypes.byref(ob.s), n, _p(orig, c_d), _p(dirs, c_d), int(any_hit),
                    _p(tri, c_i32), _p(t, c_d), _p(a, c_d), _p(be, c_d), _p(nrm, c_d), _p(visits, c_i64), nthreads)
    return tri, t, a, be, nrm, visits


def trace_brute(hm, leafpos, orig, dirs):
    om = OMesh(hm)
    n = orig.shape[0]
    tri = np.empty(n, np.int32)
    t = np.empty(n)
    lib().orc_trace_brute(ctypes.byref(om.s), _p(np.ascontiguousarray(leafpos, np.int32), c_i32), n,
                          _p(np.ascontiguousarray(orig, np.float64), c_d), _p(np.ascontiguousarray(dirs, np.float64), c_d),
                          _p(tri, c_i32), _p(t, c_d))
    return tri, t


def _mat(m):
    o = Material()
    for k in ("ka", "kd", "ks", "kt", "ke"):
        getattr(o, k)[:] = getattr(m, k)[:]
    o.shininess = m.shininess
    return o


def render(hm, b, mats, light, cam, opts, pixels=None, nthreads=0, fp64=False, records=False, branches=False):
    """mats/light/cam/opts may be rtamd ctypes structs (same layout) or oracle ones.

    Returns (rgb float32, counts); with fp64 also the image in doubles, with
    records also the (pixels, spp, depth+1, 2) path record (hit triangle,
    occluder code), with branches also the (pixels, spp, depth+1) shading
    branch bits."""
    om = OMesh(hm)
    ob = OBvh(b)
    marr = (Material * len(mats))(*[_mat(m) for m in mats])
    L = Light()
    L.position[:], L.color[:] = light.position[:], light.color[:]
    L.radius, L.c0, L.c1, L.c2 = light.radius, light.c0, light.c1, light.c2
    C = Camera(cam.width, cam.height, cam.fov_deg, cam.fstop, cam.focal_point)
    C.eye[:] = cam.eye[:]
    O = RenderOpts(opts.spp, opts.depth, opts.seed, opts.flags)
    rgb = np.zeros(cam.width * cam.height * 3, np.float32)
    counts = np.zeros(3, np.int64)
    if pixels is not None:
        px = np.ascontiguousarray(pixels, np.int64)
        pp, npix = _p(px, c_i64), px.size
    else:
        pp, npix = None, 0
    if not fp64 and not records and not branches:
        lib().orc_render(ctypes.byref(om.s), ctypes.byref(ob.s), marr, len(mats), ctypes.byref(L),
                         ctypes.byref(C), ctypes.byref(O), _p(rgb, c_f), pp, npix, _p(counts, c_i64), nthreads)
        return rgb, counts
    rgb64 = np.zeros(cam.width * cam.height * 3, np.float64)
    rec = np.full((cam.width * cam.height, opts.spp, opts.depth + 1, 2), -1, np.int32)
    br = np.full((cam.width * cam.height, opts.spp, opts.depth + 1), -1, np.int32)
    lib().orc_render64(ctypes.byref(om.s), ctypes.byref(ob.s), marr, len(mats), ctypes.byref(L), ctypes.byref(C),
                       ctypes.byref(O), _p(rgb, c_f), _p(rgb64, c_d), _p(rec, c_i32), _p(br, c_i32), pp, npix,
                       _p(counts, c_i64), nthreads)
    out = (rgb, counts)
    if fp64:
        out += (rgb64,)
    if records:
        out += (rec,)
    if branches:
        out += (br,)
    return out


def sincos(x):
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib().orc_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def cosine_dir(n, u1, u2):
    """Cosine-weighted bounce direction about the unit normal n (src/ray.cpp:5-32 frame)."""
    d = np.empty(3)
    lib().orc_cosine_dir(_p(np.ascontiguousarray(n, np.float64), c_d), float(u1), float(u2), _p(d, c_d))
    return d


def _cam(cam):
    C = Camera(cam.width, cam.height, cam.fov_deg, cam.fstop, cam.focal_point)
    C.eye[:] = cam.eye[:]
    return C


def primary_ray(cam, px, py, sample=0, seed=0):
    o = np.empty(3)
    d = np.empty(3)
    O = RenderOpts(1, 0, seed, 0)
    lib().orc_primary_ray(ctypes.byref(_cam(cam)), ctypes.byref(O), px, py, sample, _p(o, c_d), _p(d, c_d))
    return o, d


def intersect_triangle(hm, tri, orig, dirs):
    om = OMesh(hm)
    t = ctypes.c_double()
    a = ctypes.c_double()
    b = ctypes.c_double()
    nrm = np.empty(3)
    ok = lib().orc_intersect_triangle(ctypes.byref(om.s), tri, _p(np.ascontiguousarray(orig, np.float64), c_d),
                                      _p(np.ascontiguousarray(dirs, np.float64), c_d), ctypes.byref(t),
                                      ctypes.byref(a), ctypes.byref(b), _p(nrm, c_d))
    return bool(ok), t.value, a.value, b.value, nrm


def box_hit(lo, hi, orig, dirs):
    return bool(lib().orc_box_hit(_p(np.ascontiguousarray(lo, np.float32), c_f),
                                  _p(np.ascontiguousarray(hi, np.float32), c_f),
                                  _p(np.ascontiguousarray(orig, np.float64), c_d),
                                  _p(np.ascontiguousarray(dirs, np.float64), c_d)))


def wang_hash(s):
    return lib().orc_wang_hash(s)


def xorshift(state):
    st = ctypes.c_uint32(state)
    v = lib().orc_xorshift(ctypes.byref(st))
    return v, st.value


def rand_double(state):
    st = ctypes.c_uint32(state)
    v = lib().orc_rand_double(ctypes.byref(st))
    return v, st.value


def seed_index(px, py, width):
    return lib().orc_seed_index(px, py, width)


def expand_bits10(v):
    return lib().orc_expand_bits10(v)
