"""Differentiable resolve: the shading of a rendered view as a function of
material factors, light and vertex data, with its visibility held fixed.

    r = Renderer(scene, cfg); r.set_camera(cam); r.frame()   # visibility
    d = DifferentiableResolve(r)
    p = d.parameters(requires_grad=True)                       # scene values
    radiance = d(**p)                                          # [H, W, 4]
    loss = ((radiance[..., :3] - target) ** 2).sum(); loss.backward()

The autograd Function is C++ (csrc/torch_ext.cpp, torch::autograd::Function)
over lh_resolve_ex / lh_resolve_backward; `raw_backward` calls the adjoint
through the C ABI directly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .lib import check, current_stream, lib


class ResolveParams(C.Structure):
    """lh_resolve_params (include/lh/lh.h)."""

    _fields_ = [("base_color_factor", C.c_void_p), ("roughness_factor", C.c_void_p), ("metallic_factor", C.c_void_p),
                ("sun_dir", C.c_void_p), ("sun_intensity", C.c_void_p), ("frozen", C.c_void_p)]


class ResolveGrads(C.Structure):
    """lh_resolve_grads (include/lh/lh.h)."""

    _fields_ = [("base_color_factor", C.c_void_p), ("roughness_factor", C.c_void_p), ("metallic_factor", C.c_void_p),
                ("sun_dir", C.c_void_p), ("sun_intensity", C.c_void_p), ("geometry", C.c_void_p),
                ("geometry_base", C.c_uint64), ("geometry_bytes", C.c_uint64)]


_ext = None


def extension():
    """The libtorch extension (built by build_torch.py); raises if missing."""
    global _ext
    if _ext is None:
        import importlib  # noqa: PLC0415

        lib()  # the HIP library first (its tuning is applied once)
        try:
            _ext = importlib.import_module("lighthugger_amd._lh_torch")
        except ImportError as e:
            raise RuntimeError("lighthugger_amd/_lh_torch is missing: build it (__graft_entry__.build())") from e
    return _ext


def _bind():
    L = lib()
    if not getattr(L, "_diff_bound", False):
        L.lh_resolve_ex.restype = C.c_int32
        L.lh_resolve_ex.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(ResolveParams), C.c_void_p]
        L.lh_resolve_backward.restype = C.c_int32
        L.lh_resolve_backward.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(ResolveParams), C.c_void_p,
                                          C.POINTER(ResolveGrads), C.c_void_p]
        L._diff_bound = True
    return L


PARAM_NAMES = ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity")


class DifferentiableResolve:
    def __init__(self, renderer):
        import torch  # noqa: PLC0415

        self.torch = torch
        self.r = renderer
        self.geometry = renderer.scene.geometry_words()  # aliases the scene's geometry arena
        cfg = renderer.view.cfg
        # per-pixel quantities the adjoint holds fixed, recorded by forward_raw
        self.frozen = torch.empty(cfg.width * cfg.height * 4, dtype=torch.int32, device=self.geometry.device)
        self._recorded = False

    def parameters(self, requires_grad: bool = False) -> dict:
        """The differentiable parameters as the scene and the view's uniforms hold them."""
        t = self.torch
        hs = self.r.host_scene
        u = self.r.view.uniforms
        dev = self.geometry.device
        p = {
            "base_color": t.tensor(np.asarray(hs.mesh_infos["base_color_factor"], np.float32).reshape(-1, 3), device=dev),
            "roughness": t.tensor(hs.factors["roughness_factor"].astype(np.float32), device=dev),
            "metallic": t.tensor(hs.factors["metallic_factor"].astype(np.float32), device=dev),
            "sun_dir": t.tensor(np.asarray(u["sun_dir"], np.float32).ravel(), device=dev),
            "sun_intensity": t.tensor(np.asarray(u["sun_intensity"], np.float32).ravel(), device=dev),
        }
        for v in p.values():
            v.requires_grad_(requires_grad)
        return p

    def __call__(self, base_color, roughness, metallic, sun_dir, sun_intensity, geometry=None):
        g = self.geometry if geometry is None else geometry
        return extension().resolve(C.addressof(self.r.scene.c), C.addressof(self.r.view.c), current_stream(), g,
                                   self.r.scene.geometry_base, base_color.contiguous(), roughness.contiguous(),
                                   metallic.contiguous(), sun_dir.contiguous(), sun_intensity.contiguous())

    def _params_struct(self, p: dict, frozen: bool) -> ResolveParams:
        return ResolveParams(*[int(p[k].data_ptr()) for k in PARAM_NAMES],
                             int(self.frozen.data_ptr()) if frozen else 0)

    def forward_raw(self, p: dict) -> None:
        """lh_resolve_ex into the view's radiance buffer; records the adjoint's
        fixed quantities for raw_backward(..., recorded=True)."""
        s = self._params_struct(p, True)
        check(_bind().lh_resolve_ex(self.r.scene.address, self.r.view.address, C.byref(s), current_stream()),
              "lh_resolve_ex")
        self._recorded = True

    def raw_backward(self, p: dict, grad_radiance, out: dict | None = None, recorded: bool = False) -> dict:
        """Gradients of sum(grad_radiance * radiance) through lh_resolve_backward
        (added into `out` when given, else into fresh zero tensors).  With
        `recorded`, the fixed quantities come from the last forward_raw."""
        if recorded and not self._recorded:
            raise RuntimeError("raw_backward(recorded=True) needs a forward_raw first")
        t = self.torch
        if out is None:
            out = {k: t.zeros_like(p[k]) for k in PARAM_NAMES}
            out["geometry"] = t.zeros_like(self.geometry)
        g = grad_radiance.contiguous()
        assert g.dtype == t.float32 and g.numel() == self.r.view.cfg.width * self.r.view.cfg.height * 4
        grads = ResolveGrads(*[int(out[k].data_ptr()) for k in PARAM_NAMES], int(out["geometry"].data_ptr()),
                             self.r.scene.geometry_base, int(out["geometry"].numel()) * 4)
        s = self._params_struct(p, recorded)
        check(_bind().lh_resolve_backward(self.r.scene.address, self.r.view.address, C.byref(s), int(g.data_ptr()),
                                          C.byref(grads), current_stream()), "lh_resolve_backward")
        return out

    def unsupported(self) -> bool:
        """Whether the last backward skipped pixels of unsupported mesh infos."""
        return bool(self.r.status()[abi.CNT_FLAGS] & abi.FLAG_BACKWARD_UNSUPPORTED)
