"""Differentiable rendering: a torch.autograd.Function over the library's differentiable mode.

The reference's Python preprocessing package (preprocessing/splat_oracle) holds no autograd
code; this module is its host for the differentiable form.  Forward = the compositor path with
opacity logits and fp32 rasterizer inputs (HDSR_DIFFERENTIABLE), rendered straight into torch
tensors; backward = hdsr_backward (hand-written HIP: reverse walk of each pixel's list with
transmittance recovered by division, per-tile pre-reduction, then the projection's chain rule).

Parameters (float32 CUDA tensors, one row per splat):
  means (N, 3), log_scales (N, 3), quats (N, 4) as (w, x, y, z), opacity_logits (N,),
  sh (N, 3 + 3K + 3) packed [dc rgb | K SH coefficients rgb-interleaved | emissive rgb].
Output: rgba (H, W, 4) and depth (H, W) (the depth gets no gradient).

There is no CPU fallback: without the built library or a GPU this module raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, load
from .camera import Camera, frame_uniforms_for
from .compositor import _plan_struct, planTileSplatCompositor


def sh_stride(deg: int) -> int:
    return 3 + 3 * ((deg + 1) ** 2 - 1) + 3


class DiffRenderer:
    """One differentiable compositor context: N splats of SH degree d at one resolution.

    The context keeps the tile lists and per-pixel state of its last render; a backward for an
    older render of the same context re-renders it first (the forward is deterministic)."""

    def __init__(self, n: int, width: int, height: int, sh_degree: int, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("the differentiable renderer needs a HIP device")
        self._lib = load()
        self.n, self.width, self.height, self.sh_degree = int(n), int(width), int(height), int(sh_degree)
        self.device = torch.device(device)
        self.plan = planTileSplatCompositor({"viewportWidth": self.width, "viewportHeight": self.height,
                                             "tileSizePx": 16, "splatCount": self.n})
        h = C.c_void_p()
        check(self._lib.hdsr_create(C.byref(_plan_struct(self.plan)), _lib.HDSR_DIFFERENTIABLE, C.byref(h)))
        self.handle = h
        self.materials = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        self.sorted = torch.arange(self.n, dtype=torch.int32, device=self.device)
        self.generation = 0
        self._bound = None

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.hdsr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def uniforms(self, cam: Camera, splat_scale: float = 1.0) -> np.ndarray:
        return frame_uniforms_for(cam, self.plan, self.sh_degree, splat_scale)

    def _check_params(self, ps) -> None:
        shapes = [(self.n, 3), (self.n, 3), (self.n, 4), (self.n,), (self.n, sh_stride(self.sh_degree))]
        for t, s in zip(ps, shapes):
            if tuple(t.shape) != s or t.dtype != torch.float32 or t.device.type != "cuda" or not t.is_contiguous():
                raise ValueError(f"parameter must be a contiguous float32 CUDA tensor of shape {s}")

    def _bind(self, ps) -> None:
        key = tuple(int(t.data_ptr()) for t in ps)
        if key == self._bound:
            return
        means, log_scales, quats, logits, sh = ps
        s = _lib.Splats(means.data_ptr(), log_scales.data_ptr(), quats.data_ptr(), logits.data_ptr(),
                        self.materials.data_ptr(), None, 0, sh.data_ptr(), self.sorted.data_ptr(), self.n,
                        self.sh_degree, _lib.HDSR_MEM_DEVICE, sh.numel())
        check(self._lib.hdsr_bind_splats(self.handle, C.byref(s)), self.handle)
        self._bound = key

    def render(self, uniforms: np.ndarray, means, log_scales, quats, opacity_logits, sh):
        """Forward only (no autograd): returns (rgba, depth) tensors."""
        ps = (means, log_scales, quats, opacity_logits, sh)
        self._check_params(ps)
        self._bind(ps)
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        check(self._lib.hdsr_set_frame(self.handle, u.ctypes.data), self.handle)
        rgba = torch.empty(self.height, self.width, 4, dtype=torch.float32, device=self.device)
        depth = torch.empty(self.height, self.width, dtype=torch.float32, device=self.device)
        check(self._lib.hdsr_bind_outputs(self.handle, rgba.data_ptr(), depth.data_ptr()), self.handle)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        try:
            check(self._lib.hdsr_render(self.handle, C.c_void_p(stream)), self.handle)
            check(self._lib.hdsr_sync(self.handle), self.handle)
        finally:
            self._lib.hdsr_bind_outputs(self.handle, None, None)
        self.generation += 1
        return rgba, depth

    def backward(self, grad_rgba, means, log_scales, quats, opacity_logits, sh, out=None) -> tuple:
        """Gradients of sum(grad_rgba * rgba) for the last render: added into 'out' (five tensors
        shaped like the parameters) when given, else into fresh zero tensors."""
        g = grad_rgba.to(torch.float32).contiguous()
        if tuple(g.shape) != (self.height, self.width, 4) or g.device.type != "cuda":
            raise ValueError("grad_rgba must be a CUDA tensor of shape (H, W, 4)")
        ps = (means, log_scales, quats, opacity_logits, sh)
        if out is None:
            outs = tuple(torch.zeros_like(t) for t in ps)
        else:
            outs = tuple(out)
            for o, t in zip(outs, ps):
                if o.shape != t.shape or o.dtype != torch.float32 or not o.is_cuda or not o.is_contiguous():
                    raise ValueError("gradient buffers must be contiguous float32 CUDA tensors shaped like the parameters")
        gr = _lib.Grads(*[t.data_ptr() for t in outs])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(self._lib.hdsr_backward(self.handle, g.data_ptr(), C.byref(gr), C.c_void_p(stream)), self.handle)
        return outs

    def stats(self) -> dict:
        s = _lib.Stats()
        check(self._lib.hdsr_get_stats(self.handle, C.byref(s)), self.handle)
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_}

    def debug(self, name: str) -> np.ndarray:
        which = _lib.DBG[name]
        nbytes = C.c_uint64()
        check(self._lib.hdsr_debug_size(self.handle, which, C.byref(nbytes)), self.handle)
        out = np.empty(nbytes.value // 4, np.uint32)
        check(self._lib.hdsr_debug_readback(self.handle, which, out.ctypes.data, out.nbytes), self.handle)
        return out

    def read_output(self, which: int, dtype) -> np.ndarray:
        out = np.empty((self.height, self.width), dtype)
        check(self._lib.hdsr_readback(self.handle, which, out.ctypes.data, out.nbytes), self.handle)
        return out


class RenderSplats(torch.autograd.Function):
    """rgba, depth = RenderSplats.apply(renderer, uniforms, means, log_scales, quats, logits, sh)."""

    @staticmethod
    def forward(ctx, renderer: DiffRenderer, uniforms, means, log_scales, quats, opacity_logits, sh):
        rgba, depth = renderer.render(uniforms, means, log_scales, quats, opacity_logits, sh)
        ctx.renderer = renderer
        ctx.uniforms = uniforms
        ctx.generation = renderer.generation
        ctx.save_for_backward(means, log_scales, quats, opacity_logits, sh)
        ctx.mark_non_differentiable(depth)
        return rgba, depth

    @staticmethod
    def backward(ctx, grad_rgba, grad_depth):
        r: DiffRenderer = ctx.renderer
        ps = ctx.saved_tensors
        if r.generation != ctx.generation:
            r.render(ctx.uniforms, *ps)  # restore this render's lists and per-pixel state
        grads = r.backward(grad_rgba, *ps)
        return (None, None) + grads


def render_splats(renderer: DiffRenderer, cam_or_uniforms, means, log_scales, quats, opacity_logits, sh):
    """Differentiable render of one view; returns (rgba (H, W, 4), depth (H, W))."""
    u = renderer.uniforms(cam_or_uniforms) if isinstance(cam_or_uniforms, Camera) else cam_or_uniforms
    return RenderSplats.apply(renderer, u, means, log_scales, quats, opacity_logits, sh)
