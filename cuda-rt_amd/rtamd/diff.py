"""Differentiable render through the C-ABI (rt_diff_forward / rt_diff_backward).

The scene is a rendered mesh with one Lambertian material per triangle
(kd = albedo[t]) and a light whose colour is scaled by a scalar intensity.
`render(renderer, albedo, intensity, positions)` is a torch autograd
function: the forward renders on the GPU and keeps the path record, the
backward runs the adjoint kernel on it. Visibility is held fixed (no
silhouette-edge terms); vertex normals are inputs, not differentiated.
"""
import ctypes

from . import (BVH, Camera, Light, MeshView, RenderOpts, _stream_handle, _torch, lib, render_opts, vp)

RT_DIFF_MAX_DEPTH = 7


def _bind_diff(L):
    if getattr(L, "_diff_bound", False):
        return
    L.rt_diff_record_bytes.restype = ctypes.c_size_t
    L.rt_diff_record_bytes.argtypes = [ctypes.POINTER(Camera), ctypes.POINTER(RenderOpts)]
    L.rt_diff_forward.argtypes = [vp, vp, ctypes.POINTER(MeshView), vp, ctypes.c_double, ctypes.POINTER(Light),
                                  ctypes.POINTER(Camera), ctypes.POINTER(RenderOpts), vp, vp, vp]
    L.rt_diff_backward.argtypes = [vp, vp, ctypes.POINTER(MeshView), vp, ctypes.c_double, ctypes.POINTER(Light),
                                   ctypes.POINTER(Camera), ctypes.POINTER(RenderOpts), vp, vp, vp, vp, vp, vp]
    L._diff_bound = True


class DiffRenderer:
    """Scene state of the differentiable render: device mesh, its BVH (refit
    when positions change), light, camera and render options."""

    def __init__(self, ctx, mesh, light, cam, spp, depth, seed=0, morton_bits=30, flags=0):
        self.ctx, self.L = ctx, lib()
        _bind_diff(self.L)
        if depth > RT_DIFF_MAX_DEPTH:
            raise ValueError(f"depth {depth} exceeds {RT_DIFF_MAX_DEPTH}")
        self.mesh = mesh
        self.light = light
        self.cam = cam
        self.opts = render_opts(spp, depth=depth, seed=seed, flags=flags)
        self.bvh = BVH(ctx, mesh, morton_bits=morton_bits)
        torch = _torch()
        nbytes = self.L.rt_diff_record_bytes(ctypes.byref(cam), ctypes.byref(self.opts))
        self.records = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")

    @property
    def shape(self):
        return (self.cam.height, self.cam.width, 3)

    def set_positions(self, positions, stream=None):
        """positions: (nv, 3) tensor on the device; refits the BVH."""
        p = positions.detach().to(_torch().float32)
        self.mesh.set_positions(p[:, 0], p[:, 1], p[:, 2])
        self.bvh.refit(self.mesh, stream=stream)

    def forward(self, albedo, intensity, image=None, stream=None):
        torch = _torch()
        albedo = albedo.detach().to(torch.float64).contiguous()
        if albedo.shape != (self.mesh.n, 3):
            raise ValueError(f"albedo must be ({self.mesh.n}, 3)")
        if image is None:
            image = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        mv = self.mesh.view()
        self.ctx.check(self.L.rt_diff_forward(self.ctx.h, self.bvh.h, ctypes.byref(mv), albedo.data_ptr(),
                                              float(intensity), ctypes.byref(self.light), ctypes.byref(self.cam),
                                              ctypes.byref(self.opts), image.data_ptr(), self.records.data_ptr(),
                                              _stream_handle(stream)))
        return image

    def backward(self, albedo, intensity, dimage, stream=None):
        torch = _torch()
        albedo = albedo.detach().to(torch.float64).contiguous()
        dimage = dimage.detach().to(torch.float64).contiguous()
        g_alb = torch.empty((self.mesh.n, 3), dtype=torch.float64, device="cuda")
        g_light = torch.empty((), dtype=torch.float64, device="cuda")
        g_pos = torch.empty((self.mesh.nv, 3), dtype=torch.float64, device="cuda")
        mv = self.mesh.view()
        self.ctx.check(self.L.rt_diff_backward(self.ctx.h, self.bvh.h, ctypes.byref(mv), albedo.data_ptr(),
                                               float(intensity), ctypes.byref(self.light), ctypes.byref(self.cam),
                                               ctypes.byref(self.opts), self.records.data_ptr(), dimage.data_ptr(),
                                               g_alb.data_ptr(), g_light.data_ptr(), g_pos.data_ptr(),
                                               _stream_handle(stream)))
        return g_alb, g_light, g_pos


def _make_function():
    torch = _torch()

    class RenderFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, albedo, intensity, positions, renderer):
            if positions is not None:
                renderer.set_positions(positions)
            s = float(intensity)
            image = renderer.forward(albedo, s)
            ctx.renderer = renderer
            ctx.s = s
            ctx.has_pos = positions is not None
            ctx.save_for_backward(albedo)
            return image

        @staticmethod
        def backward(ctx, grad):
            (albedo,) = ctx.saved_tensors
            g_alb, g_light, g_pos = ctx.renderer.backward(albedo, ctx.s, grad)
            return (g_alb.to(albedo.dtype), g_light, g_pos if ctx.has_pos else None, None)

    return RenderFunction


_fn = None


def render(renderer, albedo, intensity, positions=None):
    """Differentiable image (H, W, 3) float64; intensity is a 0-d tensor or
    float; positions (nv, 3) or None to keep the mesh's current ones. The
    path record of this call is what the backward replays, so call backward
    before the next render with the same renderer."""
    global _fn
    torch = _torch()
    if _fn is None:
        _fn = _make_function()
    if not torch.is_tensor(intensity):
        intensity = torch.tensor(float(intensity), dtype=torch.float64, device="cuda")
    return _fn.apply(albedo, intensity, positions, renderer)
