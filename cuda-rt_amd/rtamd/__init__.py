"""Python binding of librtamd (the MI355X ray-tracer core) over its C-ABI.

The binding only marshals pointers: device memory comes from torch tensors
on the HIP device and every computation runs in librtamd's HIP kernels.
There is no CPU fallback: if the library is missing, loading it raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "librtamd.so")

RT_OK, RT_ERR_INVALID, RT_ERR_OOM, RT_ERR_DEVICE, RT_ERR_UNSUPPORTED, RT_ERR_OVERFLOW = range(6)
RT_TRACE_CLOSEST, RT_TRACE_ANY = 0, 1
RT_BUILD_DEFAULT, RT_BUILD_KARRAS_SEARCH = 0, 1
RT_RENDER_RANDVEC_RTL, RT_RENDER_CLAMP, RT_RENDER_WAVEFRONT, RT_RENDER_ONE_STREAM = 1, 2, 4, 8

c_f = ctypes.POINTER(ctypes.c_float)
c_i32 = ctypes.POINTER(ctypes.c_int32)
c_u32 = ctypes.POINTER(ctypes.c_uint32)
c_d = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


class MeshView(ctypes.Structure):
    _fields_ = [("vx", vp), ("vy", vp), ("vz", vp), ("nx", vp), ("ny", vp), ("nz", vp),
                ("idx", vp), ("mat_id", vp), ("nv", ctypes.c_int32), ("n", ctypes.c_int32)]


class BuildOpts(ctypes.Structure):
    _fields_ = [("morton_bits", ctypes.c_int32), ("flags", ctypes.c_int32)]


class ExportBufs(ctypes.Structure):
    _fields_ = [("codes", vp), ("ids", vp), ("parent", vp), ("left", vp), ("right", vp),
                ("lo", vp), ("hi", vp), ("tri_lo", vp), ("tri_hi", vp)]


class Rays(ctypes.Structure):
    _fields_ = [("orig", vp), ("dir", vp), ("n", ctypes.c_int64)]


class Hits(ctypes.Structure):
    _fields_ = [("tri", vp), ("t", vp), ("alpha", vp), ("beta", vp)]


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


EXPORTED = [
    "rt_version", "rt_ctx_create", "rt_ctx_destroy", "rt_last_error", "rt_bvh_build", "rt_bvh_rebuild",
    "rt_bvh_refit", "rt_bvh_export", "rt_bvh_scene_bounds", "rt_bvh_num_triangles", "rt_bvh_diagnostics", "rt_bvh_destroy",
    "rt_trace", "rt_render", "rt_gen_soup", "rt_gen_icosphere_counts", "rt_gen_icosphere", "rt_gen_jitter",
    "rt_ctx_set_kernel_timing", "rt_ctx_kernel_stats", "rt_render_noalloc", "rt_render_noalloc_max_bytes",
    "rt_bvh_create", "rt_bvh_compute_bounds", "rt_bvh_compute_morton", "rt_bvh_sort_morton", "rt_bvh_build_tree",
    "rt_ctx_fail_alloc", "rt_ctx_device_bytes",
]


class KernelStat(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("ms", ctypes.c_double), ("box_tests", ctypes.c_uint64),
                ("tri_tests", ctypes.c_uint64), ("rays", ctypes.c_uint64)]

_lib = None


def lib():
    """Load librtamd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: run python cuda-rt_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    L.rt_version.restype = ctypes.c_char_p
    L.rt_last_error.restype = ctypes.c_char_p
    L.rt_last_error.argtypes = [vp]
    L.rt_ctx_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    L.rt_ctx_destroy.argtypes = [vp]
    L.rt_bvh_build.argtypes = [vp, ctypes.POINTER(MeshView), ctypes.POINTER(BuildOpts), vp, ctypes.POINTER(vp)]
    L.rt_bvh_create.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(BuildOpts), ctypes.POINTER(vp)]
    L.rt_bvh_compute_bounds.argtypes = [vp, ctypes.POINTER(MeshView), vp, vp, vp]
    L.rt_bvh_compute_morton.argtypes = [vp, c_d, c_d, vp]
    L.rt_bvh_sort_morton.argtypes = [vp, vp]
    L.rt_bvh_build_tree.argtypes = [vp, ctypes.POINTER(MeshView), vp]
    L.rt_bvh_rebuild.argtypes = [vp, ctypes.POINTER(MeshView), vp]
    L.rt_bvh_refit.argtypes = [vp, ctypes.POINTER(MeshView), vp]
    L.rt_bvh_export.argtypes = [vp, ctypes.POINTER(ExportBufs), vp]
    L.rt_bvh_scene_bounds.argtypes = [vp, vp, c_d, c_d]
    L.rt_bvh_num_triangles.argtypes = [vp]
    L.rt_bvh_diagnostics.argtypes = [vp, vp, c_u32]
    L.rt_bvh_destroy.argtypes = [vp]
    L.rt_trace.argtypes = [vp, ctypes.POINTER(MeshView), ctypes.POINTER(Rays), ctypes.c_int32, ctypes.c_double,
                           ctypes.POINTER(Hits), vp]
    L.rt_render.argtypes = [vp, vp, ctypes.POINTER(MeshView), vp, ctypes.c_int32, ctypes.POINTER(Light),
                            ctypes.POINTER(Camera), ctypes.POINTER(RenderOpts), vp, vp, vp]
    L.rt_render_noalloc.argtypes = L.rt_render.argtypes
    L.rt_render_noalloc_max_bytes.restype = ctypes.c_size_t
    L.rt_gen_soup.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, c_d,
                              ctypes.c_double] + [c_f] * 6 + [c_i32]
    L.rt_gen_icosphere_counts.argtypes = [ctypes.c_int32, c_i32, c_i32]
    L.rt_gen_icosphere.argtypes = [ctypes.c_int32, ctypes.c_double, c_d, ctypes.c_double, ctypes.c_uint64] + \
        [c_f] * 6 + [c_i32]
    L.rt_gen_jitter.argtypes = [ctypes.c_int32, c_f, c_f, c_f, ctypes.c_double, ctypes.c_uint64, c_f, c_f, c_f]
    L.rt_ctx_set_kernel_timing.argtypes = [vp, ctypes.c_int32]
    L.rt_ctx_fail_alloc.argtypes = [vp, ctypes.c_int32]
    L.rt_ctx_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.rt_ctx_kernel_stats.argtypes = [vp, ctypes.POINTER(KernelStat)]
    _lib = L
    _bind_batch(L)
    return L


def _p(a, ct):
    return a.ctypes.data_as(ct)


# ------------------------------------------------------------------ host scenes
class HostMesh:
    """fp32 SoA mesh in host memory (numpy)."""

    def __init__(self, vx, vy, vz, nx, ny, nz, idx, mat_id=None):
        self.vx, self.vy, self.vz = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
        self.nx, self.ny, self.nz = [np.ascontiguousarray(a, np.float32) for a in (nx, ny, nz)]
        self.idx = np.ascontiguousarray(idx, np.int32).reshape(-1)
        self.mat_id = None if mat_id is None else np.ascontiguousarray(mat_id, np.int32)

    @property
    def n(self):
        return self.idx.size // 3

    @property
    def nv(self):
        return self.vx.size

    @staticmethod
    def concat(meshes):
        off = 0
        parts = {k: [] for k in ("vx", "vy", "vz", "nx", "ny", "nz", "idx", "mat")}
        for mi, m in enumerate(meshes):
            for k in ("vx", "vy", "vz", "nx", "ny", "nz"):
                parts[k].append(getattr(m, k))
            parts["idx"].append(m.idx + off)
            parts["mat"].append(m.mat_id if m.mat_id is not None else np.full(m.n, mi, np.int32))
            off += m.nv
        return HostMesh(*[np.concatenate(parts[k]) for k in ("vx", "vy", "vz", "nx", "ny", "nz", "idx")],
                        mat_id=np.concatenate(parts["mat"]))


def gen_soup(n, seed, edge_median=0.002, edge_sigma=0.5, offset=(-0.5, -0.5, -2.0), scale=1.0):
    L = lib()
    nv = 3 * n
    a = [np.empty(nv, np.float32) for _ in range(6)]
    idx = np.empty(3 * n, np.int32)
    off = np.asarray(offset, np.float64)
    L.rt_gen_soup(n, seed, edge_median, edge_sigma, _p(off, c_d), scale, *[_p(x, c_f) for x in a], _p(idx, c_i32))
    return HostMesh(*a, idx)


def gen_icosphere(freq, radius=1.0, center=(0.0, 0.0, 0.0), noise_amp=0.0, noise_seed=0):
    L = lib()
    nv, n = ctypes.c_int32(), ctypes.c_int32()
    L.rt_gen_icosphere_counts(freq, ctypes.byref(nv), ctypes.byref(n))
    a = [np.empty(nv.value, np.float32) for _ in range(6)]
    idx = np.empty(3 * n.value, np.int32)
    c = np.asarray(center, np.float64)
    L.rt_gen_icosphere(freq, radius, _p(c, c_d), noise_amp, noise_seed, *[_p(x, c_f) for x in a], _p(idx, c_i32))
    return HostMesh(*a, idx)


def gen_jitter(m, sigma, seed):
    L = lib()
    out = [np.empty(m.nv, np.float32) for _ in range(3)]
    L.rt_gen_jitter(m.nv, _p(m.vx, c_f), _p(m.vy, c_f), _p(m.vz, c_f), sigma, seed, *[_p(x, c_f) for x in out])
    return HostMesh(out[0], out[1], out[2], m.nx, m.ny, m.nz, m.idx, m.mat_id)


def ground_quad(x0=-4.0, x1=4.0, y=-1.0, z0=-8.0, z1=-0.5):
    vx = np.array([x0, x1, x1, x0], np.float32)
    vy = np.full(4, y, np.float32)
    vz = np.array([z1, z1, z0, z0], np.float32)
    nx = np.zeros(4, np.float32)
    ny = np.ones(4, np.float32)
    nz = np.zeros(4, np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.int32)
    return HostMesh(vx, vy, vz, nx, ny, nz, idx)


def material(kd=(0.8, 0.8, 0.8), ks=(0, 0, 0), kt=(0, 0, 0), ke=(0, 0, 0), ka=(0, 0, 0), shininess=0.0):
    m = Material()
    m.ka[:], m.kd[:], m.ks[:], m.kt[:], m.ke[:] = ka, kd, ks, kt, ke
    m.shininess = shininess
    return m


def light(position=(1.0, 1.0, 1.0), color=(1.0, 1.0, 1.0), radius=0.5, c0=0.0, c1=0.01, c2=0.001):
    """Defaults follow Light_h's constructor (inc/Light.h:28)."""
    L = Light()
    L.position[:], L.color[:] = position, color
    L.radius, L.c0, L.c1, L.c2 = radius, c0, c1, c2
    return L


def camera(width, height, fov_deg=30.0, fstop=8.0, focal_point=10.0, eye=(0.0, 0.0, 0.0)):
    """Defaults follow Camera::Camera (kernels/camera.cu:9-22)."""
    c = Camera()
    c.width, c.height, c.fov_deg, c.fstop, c.focal_point = width, height, fov_deg, fstop, focal_point
    c.eye[:] = eye
    return c


def render_opts(spp, depth=0, seed=0, flags=RT_RENDER_CLAMP):
    o = RenderOpts()
    o.spp, o.depth, o.seed, o.flags = spp, depth, seed, flags
    return o


# ------------------------------------------------------------------ device side
def _torch():
    import torch
    return torch


def _stream_handle(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class DeviceMesh:
    """A HostMesh copied into device tensors; keeps the tensors alive."""

    def __init__(self, hm, device="cuda"):
        torch = _torch()
        t = lambda a: torch.from_numpy(a).to(device)
        self.host = hm
        self.vx, self.vy, self.vz = t(hm.vx), t(hm.vy), t(hm.vz)
        self.nx, self.ny, self.nz = t(hm.nx), t(hm.ny), t(hm.nz)
        self.idx = t(hm.idx)
        self.mat_id = t(hm.mat_id) if hm.mat_id is not None else None
        self.n, self.nv = hm.n, hm.nv

    def set_positions(self, vx, vy, vz):
        self.vx.copy_(vx)
        self.vy.copy_(vy)
        self.vz.copy_(vz)

    def view(self):
        v = MeshView()
        v.vx, v.vy, v.vz = self.vx.data_ptr(), self.vy.data_ptr(), self.vz.data_ptr()
        v.nx, v.ny, v.nz = self.nx.data_ptr(), self.ny.data_ptr(), self.nz.data_ptr()
        v.idx = self.idx.data_ptr()
        v.mat_id = self.mat_id.data_ptr() if self.mat_id is not None else None
        v.nv, v.n = self.nv, self.n
        return v


class RtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rt status {status}: {msg}")
        self.status = status


class Context:
    def __init__(self, device=0):
        self.L = lib()
        self.h = vp()
        st = self.L.rt_ctx_create(device, ctypes.byref(self.h))
        if st != RT_OK:
            raise RtError(st, "rt_ctx_create failed")

    def check(self, st):
        if st != RT_OK:
            raise RtError(st, self.L.rt_last_error(self.h).decode())

    def fail_alloc(self, k):
        """Make the k-th device allocation from now on fail (k = 0 disarms; rt_ctx_fail_alloc)."""
        self.check(self.L.rt_ctx_fail_alloc(self.h, int(k)))

    def get_tuning(self, ntris, nrays):
        """(tuning, found) for that size bucket (rt_ctx_get_tuning)."""
        t, found = Tuning(), ctypes.c_int32()
        self.check(self.L.rt_ctx_get_tuning(self.h, int(ntris), int(nrays), ctypes.byref(t), ctypes.byref(found)))
        return t, bool(found.value)

    def set_tuning(self, ntris, nrays, t):
        """Set (or with t=None remove) the entry for that size bucket (rt_ctx_set_tuning)."""
        self.check(self.L.rt_ctx_set_tuning(self.h, int(ntris), int(nrays), None if t is None else ctypes.byref(t)))

    def save_tuning(self, path):
        self.check(self.L.rt_ctx_save_tuning(self.h, os.fsencode(path)))

    def load_tuning(self, path):
        self.check(self.L.rt_ctx_load_tuning(self.h, os.fsencode(path)))

    def device_bytes(self):
        """Device bytes this context's objects hold (rt_ctx_device_bytes)."""
        out = ctypes.c_uint64()
        self.check(self.L.rt_ctx_device_bytes(self.h, ctypes.byref(out)))
        return int(out.value)

    def set_kernel_timing(self, enable):
        """Bracket each wavefront traversal launch with HIP events (rt_ctx_set_kernel_timing)."""
        self.check(self.L.rt_ctx_set_kernel_timing(self.h, 1 if enable else 0))

    def kernel_stats(self):
        """Totals since the last call: {'closest': {...}, 'any': {...}} (rt_ctx_kernel_stats)."""
        out = (KernelStat * 2)()
        self.check(self.L.rt_ctx_kernel_stats(self.h, out))
        return {k: {"launches": int(o.launches), "ms": float(o.ms), "box_tests": int(o.box_tests),
                    "tri_tests": int(o.tri_tests), "rays": int(o.rays)} for k, o in zip(("closest", "any"), out)}

    def close(self):
        if self.h:
            self.L.rt_ctx_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BVH:
    def __init__(self, ctx, mesh, morton_bits=30, stream=None, flags=RT_BUILD_DEFAULT, staged_n=None):
        """Build the BVH of `mesh` (rt_bvh_build), or, with mesh=None and
        staged_n=n, allocate an unbuilt BVH for n triangles (rt_bvh_create)
        to build stage by stage."""
        self.ctx, self.L = ctx, ctx.L
        self.h = vp()
        self.n = mesh.n if mesh is not None else int(staged_n)
        self.bits = morton_bits
        o = BuildOpts(morton_bits, flags)
        if mesh is None:
            ctx.check(self.L.rt_bvh_create(ctx.h, self.n, ctypes.byref(o), ctypes.byref(self.h)))
            return
        mv = mesh.view()
        ctx.check(self.L.rt_bvh_build(ctx.h, ctypes.byref(mv), ctypes.byref(o), _stream_handle(stream),
                                      ctypes.byref(self.h)))

    @classmethod
    def create(cls, ctx, n, morton_bits=30, flags=RT_BUILD_DEFAULT):
        """An unbuilt BVH for n triangles (rt_bvh_create)."""
        return cls(ctx, None, morton_bits, flags=flags, staged_n=n)

    def compute_bounds(self, mesh, tri_lo=None, tri_hi=None, stream=None):
        """Triangle boxes and scene bounds (rt_bvh_compute_bounds); tri_lo /
        tri_hi: optional float32 CUDA tensors of n*3 that receive the boxes."""
        mv = mesh.view()
        ptr = lambda t: None if t is None else t.data_ptr()
        self.ctx.check(self.L.rt_bvh_compute_bounds(self.h, ctypes.byref(mv), ptr(tri_lo), ptr(tri_hi),
                                                    _stream_handle(stream)))

    def compute_morton(self, mn=None, mx=None, stream=None):
        """Morton keys normalised by mn / mx (3 doubles each), or by the scene
        bounds when both are None (rt_bvh_compute_morton)."""
        a = None if mn is None else (ctypes.c_double * 3)(*[float(v) for v in mn])
        b = None if mx is None else (ctypes.c_double * 3)(*[float(v) for v in mx])
        self.ctx.check(self.L.rt_bvh_compute_morton(self.h, a, b, _stream_handle(stream)))

    def sort_morton(self, stream=None):
        self.ctx.check(self.L.rt_bvh_sort_morton(self.h, _stream_handle(stream)))

    def build_tree(self, mesh, stream=None):
        mv = mesh.view()
        self.ctx.check(self.L.rt_bvh_build_tree(self.h, ctypes.byref(mv), _stream_handle(stream)))

    def rebuild(self, mesh, stream=None):
        mv = mesh.view()
        self.ctx.check(self.L.rt_bvh_rebuild(self.h, ctypes.byref(mv), _stream_handle(stream)))

    def refit(self, mesh, stream=None):
        mv = mesh.view()
        self.ctx.check(self.L.rt_bvh_refit(self.h, ctypes.byref(mv), _stream_handle(stream)))

    def scene_bounds(self, stream=None):
        mn = (ctypes.c_double * 3)()
        mx = (ctypes.c_double * 3)()
        self.ctx.check(self.L.rt_bvh_scene_bounds(self.h, _stream_handle(stream), mn, mx))
        return np.array(mn[:]), np.array(mx[:])

    def diagnostics(self, stream=None):
        """(traversal stack overflows, build consistency failures); both 0 when healthy."""
        out = (ctypes.c_uint32 * 2)()
        self.ctx.check(self.L.rt_bvh_diagnostics(self.h, _stream_handle(stream), out))
        return int(out[0]), int(out[1])

    def export(self, stream=None):
        """Copy the BVH into the reference layout; returns host numpy arrays."""
        torch = _torch()
        n = self.n
        nn = 2 * n - 1
        dev = torch.device("cuda")
        kd = torch.int32 if self.bits == 30 else torch.int64
        out = {
            "codes": torch.empty(n, dtype=kd, device=dev),
            "ids": torch.empty(n, dtype=torch.int32, device=dev),
            "parent": torch.empty(nn, dtype=torch.int32, device=dev),
            "left": torch.empty(nn, dtype=torch.int32, device=dev),
            "right": torch.empty(nn, dtype=torch.int32, device=dev),
            "lo": torch.empty(nn * 3, dtype=torch.float32, device=dev),
            "hi": torch.empty(nn * 3, dtype=torch.float32, device=dev),
            "tri_lo": torch.empty(n * 3, dtype=torch.float32, device=dev),
            "tri_hi": torch.empty(n * 3, dtype=torch.float32, device=dev),
        }
        e = ExportBufs(*[out[k].data_ptr() for k in ("codes", "ids", "parent", "left", "right", "lo", "hi",
                                                      "tri_lo", "tri_hi")])
        self.ctx.check(self.L.rt_bvh_export(self.h, ctypes.byref(e), _stream_handle(stream)))
        torch.cuda.synchronize()
        res = {k: v.cpu().numpy() for k, v in out.items()}
        res["codes"] = res["codes"].view(np.uint32 if self.bits == 30 else np.uint64)
        res["ids"] = res["ids"].view(np.uint32)
        for k in ("lo", "hi", "tri_lo", "tri_hi"):
            res[k] = res[k].reshape(-1, 3)
        return res

    def trace(self, mesh, orig, dirs, mode=RT_TRACE_CLOSEST, tmin=0.0, stream=None):
        """orig/dirs: (n,3) float64 device tensors. Returns (tri, t, alpha, beta) device tensors."""
        torch = _torch()
        n = orig.shape[0]
        tri = torch.empty(n, dtype=torch.int32, device=orig.device)
        t = torch.empty(n, dtype=torch.float64, device=orig.device)
        a = torch.empty_like(t)
        b = torch.empty_like(t)
        rays = Rays(orig.data_ptr(), dirs.data_ptr(), n)
        hits = Hits(tri.data_ptr(), t.data_ptr(), a.data_ptr(), b.data_ptr())
        mv = mesh.view()
        self.ctx.check(self.L.rt_trace(self.h, ctypes.byref(mv), ctypes.byref(rays), mode, tmin, ctypes.byref(hits),
                                       _stream_handle(stream)))
        return tri, t, a, b

    def close(self):
        if self.h:
            self.L.rt_bvh_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def materials_tensor(mats, device="cuda"):
    torch = _torch()
    raw = b"".join(bytes(m) for m in mats)
    arr = np.frombuffer(raw, dtype=np.uint8).copy()
    return torch.from_numpy(arr).to(device)


def render(ctx, bvh, mesh, mats_dev, nmat, light_, cam, opts, rgb=None, counters=None, stream=None):
    torch = _torch()
    if rgb is None:
        rgb = torch.empty(cam.height * cam.width * 3, dtype=torch.float32, device="cuda")
    mv = mesh.view()
    ctx.check(ctx.L.rt_render(ctx.h, bvh.h, ctypes.byref(mv), mats_dev.data_ptr(), nmat, ctypes.byref(light_),
                              ctypes.byref(cam), ctypes.byref(opts), rgb.data_ptr(),
                              counters.data_ptr() if counters is not None else None, _stream_handle(stream)))
    return rgb


def render_autotune(ctx, bvh, mesh, mats_dev, nmat, light_, cam, opts, rgb=None, stream=None):
    """rt_render_autotune: the fastest tuning for this scene's size, stored in
    the context's table; returns (tuning, ms, image of the last trial)."""
    torch = _torch()
    if rgb is None:
        rgb = torch.empty(cam.height * cam.width * 3, dtype=torch.float32, device="cuda")
    _bind_batch(ctx.L)
    mv = mesh.view()
    best, ms = Tuning(), ctypes.c_float()
    ctx.check(ctx.L.rt_render_autotune(ctx.h, bvh.h, ctypes.byref(mv), mats_dev.data_ptr(), nmat,
                                       ctypes.byref(light_), ctypes.byref(cam), ctypes.byref(opts), rgb.data_ptr(),
                                       _stream_handle(stream), ctypes.byref(best), ctypes.byref(ms)))
    return best, float(ms.value), rgb


def render_noalloc(ctx, bvh, mesh, mats_dev, nmat, light_, cam, opts, rgb, counters=None, stream=None):
    """rt_render_noalloc: the wavefront render on one stream with the library's
    static scratch (no device allocation; rgb must be given)."""
    mv = mesh.view()
    ctx.check(ctx.L.rt_render_noalloc(ctx.h, bvh.h, ctypes.byref(mv), mats_dev.data_ptr(), nmat,
                                      ctypes.byref(light_), ctypes.byref(cam), ctypes.byref(opts), rgb.data_ptr(),
                                      counters.data_ptr() if counters is not None else None,
                                      _stream_handle(stream)))
    return rgb


# ------------------------------------------------------------------ batches
class Tuning(ctypes.Structure):
    _fields_ = [("stack_ring", ctypes.c_int32), ("fetch_chunk", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32), ("node_layout", ctypes.c_int32), ("wg_threads", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(ctypes.Structure):
    _fields_ = [("bvh", vp), ("mesh", MeshView), ("mats", vp), ("nmat", ctypes.c_int32), ("light", Light),
                ("cam", Camera), ("opts", RenderOpts), ("rgb", vp)]


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return
    L.rt_batch_create.argtypes = [vp, ctypes.POINTER(BatchItem), ctypes.c_int32, ctypes.POINTER(vp)]
    L.rt_batch_rebuild.argtypes = [vp, vp]
    L.rt_batch_render.argtypes = [vp, vp, vp]
    L.rt_batch_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.rt_batch_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.rt_batch_autotune.argtypes = [vp, vp, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_float)]
    L.rt_render_autotune.argtypes = [vp, vp, ctypes.POINTER(MeshView), vp, ctypes.c_int32, ctypes.POINTER(Light),
                                     ctypes.POINTER(Camera), ctypes.POINTER(RenderOpts), vp, vp,
                                     ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_float)]
    L.rt_ctx_get_tuning.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Tuning),
                                    ctypes.POINTER(ctypes.c_int32)]
    L.rt_ctx_set_tuning.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Tuning)]
    L.rt_ctx_save_tuning.argtypes = [vp, ctypes.c_char_p]
    L.rt_ctx_load_tuning.argtypes = [vp, ctypes.c_char_p]
    L.rt_batch_diagnostics.argtypes = [vp, vp, c_u32]
    L.rt_batch_destroy.argtypes = [vp]
    L._batch_bound = True


EXPORTED += ["rt_diff_record_bytes", "rt_diff_forward", "rt_diff_backward", "rt_resolve_supersample"]
RT_RESOLVE_REFERENCE_DIVISOR = 1


def resolve_supersample(ctx, big, width, height, ss, flags=0, out=None, stream=None):
    """AverageSuperSamplingKernel (kernels/scene.cu:127-155) on a device image."""
    torch = _torch()
    L = ctx.L
    L.rt_resolve_supersample.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp]
    if out is None:
        out = torch.empty(width * height * 3, dtype=torch.float32, device="cuda")
    ctx.check(L.rt_resolve_supersample(ctx.h, big.data_ptr(), width, height, ss, flags, out.data_ptr(),
                                       _stream_handle(stream)))
    return out
EXPORTED += ["rt_obj_load", "rt_obj_counts", "rt_obj_mesh", "rt_obj_positions64", "rt_obj_materials", "rt_obj_free",
             "rt_write_bmp"]


def _bind_io(L):
    if getattr(L, "_io_bound", False):
        return
    L.rt_obj_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
    L.rt_obj_counts.argtypes = [vp] + [c_i32] * 6
    L.rt_obj_mesh.argtypes = [vp] + [c_f] * 6 + [c_i32, c_i32]
    L.rt_obj_positions64.argtypes = [vp, c_d]
    L.rt_obj_materials.argtypes = [vp, vp]
    L.rt_obj_free.argtypes = [vp]
    L.rt_write_bmp.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, c_f]
    L._io_bound = True


def load_obj(path, mtl_basepath=None):
    """Scene_h::LoadObj (src/scene.cpp:6-58): returns (HostMesh, [Material], info)."""
    L = lib()
    _bind_io(L)
    if mtl_basepath is None:
        mtl_basepath = os.path.dirname(os.path.abspath(path)) + os.sep
    h = vp()
    err = ctypes.create_string_buffer(512)
    st = L.rt_obj_load(path.encode(), mtl_basepath.encode(), ctypes.byref(h), err, len(err))
    if st != RT_OK:
        raise RtError(st, err.value.decode())
    try:
        c = [ctypes.c_int32() for _ in range(6)]
        L.rt_obj_counts(h, *[ctypes.byref(x) for x in c])
        nv, n, nmat, nshapes, raw_v, raw_vn = (x.value for x in c)
        arr = [np.empty(nv, np.float32) for _ in range(6)]
        idx = np.empty(3 * n, np.int32)
        mat_id = np.empty(n, np.int32)
        L.rt_obj_mesh(h, *[_p(a, c_f) for a in arr], _p(idx, c_i32), _p(mat_id, c_i32))
        pos64 = np.empty(3 * nv, np.float64)
        L.rt_obj_positions64(h, _p(pos64, c_d))
        mats = (Material * max(nmat, 1))()
        L.rt_obj_materials(h, mats)
    finally:
        L.rt_obj_free(h)
    hm = HostMesh(*arr, idx, mat_id)
    info = {"shapes": nshapes, "raw_vertices": raw_v, "raw_normals": raw_vn, "positions64": pos64.reshape(-1, 3)}
    return hm, [mats[k] for k in range(nmat)], info


def write_bmp(path, width, height, rgb):
    """24-bit BMP of a rendered image (writeBMP, src/bitmap.cpp:100-155)."""
    L = lib()
    _bind_io(L)
    a = np.ascontiguousarray(rgb, np.float32).reshape(-1)
    if a.size != width * height * 3:
        raise ValueError("image size mismatch")
    st = L.rt_write_bmp(path.encode(), width, height, _p(a, c_f))
    if st != RT_OK:
        raise RtError(st, f"cannot write {path}")
EXPORTED += ["rt_batch_create", "rt_batch_rebuild", "rt_batch_render", "rt_batch_set_tuning", "rt_batch_get_tuning",
             "rt_batch_autotune", "rt_batch_diagnostics", "rt_batch_destroy", "rt_render_autotune",
             "rt_ctx_get_tuning", "rt_ctx_set_tuning", "rt_ctx_save_tuning", "rt_ctx_load_tuning"]


class Batch:
    """Many scenes rendered by one set of wavefront launches (rt_batch_*).

    scenes: list of dicts with keys mesh (DeviceMesh), bvh (BVH), mats (device
    tensor), nmat, light, cam, opts, rgb (device float tensor)."""

    def __init__(self, ctx, scenes):
        self.ctx, self.L = ctx, ctx.L
        _bind_batch(self.L)
        self.scenes = scenes
        arr = (BatchItem * len(scenes))()
        for k, sc in enumerate(scenes):
            it = arr[k]
            it.bvh = sc["bvh"].h
            it.mesh = sc["mesh"].view()
            it.mats = sc["mats"].data_ptr()
            it.nmat = sc["nmat"]
            it.light = sc["light"]
            it.cam = sc["cam"]
            it.opts = sc["opts"]
            it.rgb = sc["rgb"].data_ptr()
        self._items = arr
        self.h = vp()
        ctx.check(self.L.rt_batch_create(ctx.h, arr, len(scenes), ctypes.byref(self.h)))

    def rebuild(self, stream=None):
        self.ctx.check(self.L.rt_batch_rebuild(self.h, _stream_handle(stream)))

    def render(self, counters=None, stream=None):
        self.ctx.check(self.L.rt_batch_render(self.h, counters.data_ptr() if counters is not None else None,
                                              _stream_handle(stream)))

    def set_tuning(self, t):
        self.ctx.check(self.L.rt_batch_set_tuning(self.h, ctypes.byref(t)))

    def tuning(self):
        t = Tuning()
        self.ctx.check(self.L.rt_batch_get_tuning(self.h, ctypes.byref(t)))
        return t

    def autotune(self, stream=None):
        t = Tuning()
        ms = ctypes.c_float()
        self.ctx.check(self.L.rt_batch_autotune(self.h, _stream_handle(stream), ctypes.byref(t), ctypes.byref(ms)))
        return t, ms.value

    def diagnostics(self, stream=None):
        out = (ctypes.c_uint32 * 1)()
        self.ctx.check(self.L.rt_batch_diagnostics(self.h, _stream_handle(stream), out))
        return int(out[0])

    def close(self):
        if self.h:
            self.L.rt_batch_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
