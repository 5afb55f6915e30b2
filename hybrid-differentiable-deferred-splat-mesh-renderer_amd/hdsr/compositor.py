"""Python facade of the reference's gpuTileSplatCompositor module over the C ABI.

Same names and argument meaning as src/gpuTileSplatCompositor.ts:
  TILE_SPLAT_FRAME_UNIFORM_BYTES                       (:18)
  planTileSplatCompositor(input) -> plan               (:109-139)
  createTileSplatCompositor(device, plan, options)     (:141-728)   -> Resources (destroy())
  writeTileSplatFrameUniforms(target, viewProj, plan, splatScale, shDegree, cameraPos, viewMatrix,
                              projMatrix, normalMatrix)                (:730-774)
  createTileSplatBindGroups(device, resources, splatBuffers, ...)      (:792-995)
  encodeFullComputeCompositorPipeline(encoder, resources, bindGroups)  (:1023-1179)
  encodeCompositeOnly(encoder, resources, bindGroups)                  (:1001-1014)
GPU objects become numpy arrays (host) or device pointers; the 'device' and 'encoder' arguments
are accepted for call-shape compatibility (an encoder may be a torch stream or a raw stream
handle).  Errors raise HdsrError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .camera import write_frame_uniforms
from ._lib import HdsrError, check, load

TILE_SPLAT_FRAME_UNIFORM_BYTES = 256


def planTileSplatCompositor(input: dict) -> dict:
    pin = _lib.PlanIn(int(input["viewportWidth"]), int(input["viewportHeight"]), int(input["tileSizePx"]),
                      int(input["splatCount"]), float(input.get("tileEntryMultiplier") or 0.0))
    p = _lib.Plan()
    check(load().hdsr_plan_compute(C.byref(pin), C.byref(p)))
    return {
        "viewportWidth": p.viewport_width, "viewportHeight": p.viewport_height, "tileSizePx": p.tile_size_px,
        "tileColumns": p.tile_columns, "tileRows": p.tile_rows, "tileCount": p.tile_count,
        "mortonTileCount": p.morton_tile_count, "splatCount": p.splat_count,
        "maxTotalTileRefs": p.max_total_tile_refs,
    }


def _plan_struct(plan: dict) -> _lib.Plan:
    return _lib.Plan(plan["viewportWidth"], plan["viewportHeight"], plan["tileSizePx"], plan["tileColumns"],
                     plan["tileRows"], plan["tileCount"], plan["mortonTileCount"], plan["splatCount"],
                     plan["maxTotalTileRefs"])


def _stream_handle(encoder) -> int | None:
    if encoder is None:
        return None
    if isinstance(encoder, int):
        return encoder
    h = getattr(encoder, "cuda_stream", None)
    if h is not None:
        return int(h)
    raise TypeError("encoder must be None, a raw stream handle or a torch stream")


class Resources:
    """The compositor context (createTileSplatCompositor's result)."""

    def __init__(self, plan: dict, flags: int):
        self.plan = plan
        self.flags = flags
        self._lib = load()
        h = C.c_void_p()
        rc = self._lib.hdsr_create(C.byref(_plan_struct(plan)), flags, C.byref(h))
        if rc != 0:
            raise HdsrError(rc, "hdsr_create failed: " + (self._lib.hdsr_last_error(None) or b"").decode())
        self.handle = h
        self._keep = []

    def destroy(self) -> None:
        if self.handle:
            self._lib.hdsr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- helpers beyond the reference API (readback and diagnostics) --
    def sync(self) -> None:
        check(self._lib.hdsr_sync(self.handle), self.handle)

    def read_rgba(self) -> np.ndarray:
        W, H = self.plan["viewportWidth"], self.plan["viewportHeight"]
        out = np.empty((H, W, 4), np.float32)
        check(self._lib.hdsr_readback(self.handle, _lib.OUT_RGBA, out.ctypes.data, out.nbytes), self.handle)
        return out

    def read_depth(self) -> np.ndarray:
        W, H = self.plan["viewportWidth"], self.plan["viewportHeight"]
        out = np.empty((H, W), np.float32)
        check(self._lib.hdsr_readback(self.handle, _lib.OUT_DEPTH, out.ctypes.data, out.nbytes), self.handle)
        return out

    def read_output(self, which: int, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        check(self._lib.hdsr_readback(self.handle, which, out.ctypes.data, out.nbytes), self.handle)
        return out

    def normal_params(self, **overrides) -> "_lib.NormalParams":
        """The reference's deferred-normal parameters (hdsr_normal_params_default), with overrides."""
        p = _lib.NormalParams()
        self._lib.hdsr_normal_params_default(C.byref(p))
        for k, v in overrides.items():
            setattr(p, k, v)
        return p

    def encodeScreenSpaceNormals(self, stream=None, params=None) -> None:
        """createScreenSpaceNormalsPass().encode (src/splatRenderer.ts:317-376, :1048-1057):
        G-buffer depth + normals -> the filtered-normal buffer (needs gbuffer=True)."""
        p = params if params is not None else self.normal_params()
        check(self._lib.hdsr_screen_space_normals(self.handle, C.byref(p), C.c_void_p(stream)), self.handle)

    def encodeBilateralNormalFilter(self, stream=None, params=None) -> None:
        """createBilateralNormalFilter().encode (src/splatRenderer.ts:382-442, :1063-1069): filtered
        normals + depth -> the G-buffer normal words."""
        p = params if params is not None else self.normal_params()
        check(self._lib.hdsr_bilateral_normal_filter(self.handle, C.byref(p), C.c_void_p(stream)), self.handle)

    def lighting_params(self, **overrides) -> "_lib.LightingParams":
        """The deferred-lighting parameters (hdsr_lighting_params_default: the reference viewer's
        defaults), with overrides; curves as (contrast, brightness, gamma) tuples, ao / bloom as
        device pointers (int) or None, light_direction / lighting_view_proj /
        lighting_camera_position as sequences."""
        p = _lib.LightingParams()
        self._lib.hdsr_lighting_params_default(C.byref(p))
        keep = []
        for k, v in overrides.items():
            if k == "light_direction":
                p.light_direction[:] = [float(x) for x in v]
            elif k.endswith("_curve"):
                setattr(p, k, _lib.MaterialCurve(*[float(x) for x in v]))
            elif k in ("lighting_view_proj", "lighting_camera_position"):
                a = None if v is None else np.ascontiguousarray(v, np.float32)
                keep.append(a)
                setattr(p, k, None if a is None else a.ctypes.data)
            else:
                setattr(p, k, v)
        p._keep = keep  # host arrays referenced by the struct
        return p

    def encodeDeferredLighting(self, stream=None, params=None) -> None:
        """createDeferredLightingPass().encode (src/splatRenderer.ts:446-548, :1115-1139): colour,
        depth, G-buffer normals and materials -> the lit rgba16float buffer (OUT_LIT; gbuffer=True)."""
        p = params if params is not None else self.lighting_params()
        check(self._lib.hdsr_deferred_lighting(self.handle, C.byref(p), C.c_void_p(stream)), self.handle)

    def debug(self, name: str) -> np.ndarray:
        which = _lib.DBG[name]
        n = C.c_uint64()
        check(self._lib.hdsr_debug_size(self.handle, which, C.byref(n)), self.handle)
        out = np.empty(n.value // 4, np.uint32)
        check(self._lib.hdsr_debug_readback(self.handle, which, out.ctypes.data, out.nbytes), self.handle)
        return out

    def stats(self) -> dict:
        s = _lib.Stats()
        check(self._lib.hdsr_get_stats(self.handle, C.byref(s)), self.handle)
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_ if not name.startswith("reserved")}

    def stage_times(self) -> tuple[dict, int]:
        """Mean GPU milliseconds per render of each pipeline stage (needs profile=True)."""
        ms = (C.c_float * len(_lib.STAGES))()
        n = C.c_uint32()
        check(self._lib.hdsr_stage_times(self.handle, ms, C.byref(n)), self.handle)
        return dict(zip(_lib.STAGES, [float(x) for x in ms])), int(n.value)

    def outputs(self) -> dict:
        o = _lib.Outputs()
        check(self._lib.hdsr_get_outputs(self.handle, C.byref(o)), self.handle)
        return {"rgba": o.rgba, "depth": o.depth, "gbuffer_normal": o.gbuffer_normal,
                "gbuffer_material": o.gbuffer_material, "rgba16": o.rgba16, "width": o.width, "height": o.height}

    def set_frame(self, uniforms: np.ndarray) -> None:
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        if u.nbytes != TILE_SPLAT_FRAME_UNIFORM_BYTES:
            raise ValueError("frame uniforms must be 256 bytes")
        check(self._lib.hdsr_set_frame(self.handle, u.ctypes.data), self.handle)

    def update_sorted_indices(self, sorted_indices) -> None:
        if _numel(sorted_indices) != self.plan["splatCount"]:
            raise ValueError(f"sortedIndices holds {_numel(sorted_indices)} elements; splatCount is "
                             f"{self.plan['splatCount']}")
        if hasattr(sorted_indices, "data_ptr"):
            check(self._lib.hdsr_update_sorted_indices(self.handle, sorted_indices.data_ptr(), _lib.HDSR_MEM_DEVICE),
                  self.handle)
            self._keep.append(sorted_indices)
        else:
            a = np.ascontiguousarray(sorted_indices, dtype=np.uint32)
            check(self._lib.hdsr_update_sorted_indices(self.handle, a.ctypes.data, _lib.HDSR_MEM_HOST), self.handle)


def createTileSplatCompositor(device, plan: dict, options: dict | None = None, *, gbuffer=False, debug=False,
                              rgba16=False, profile=False, bucket_scatter=False) -> Resources:
    options = options or {}
    flags = 0
    if options.get("f16"):
        flags |= _lib.HDSR_F16_COMPOSITE
    if options.get("refCaps"):
        flags |= _lib.HDSR_REF_CAPS
    if gbuffer or options.get("gbuffer"):
        flags |= _lib.HDSR_GBUFFER
    if debug or options.get("debug"):
        flags |= _lib.HDSR_DEBUG
    if rgba16 or options.get("rgba16"):
        flags |= _lib.HDSR_OUT_RGBA16
    if profile or options.get("profile"):
        flags |= _lib.HDSR_PROFILE
    if bucket_scatter or options.get("bucketScatter"):
        flags |= _lib.HDSR_BUCKET_SCATTER
    return Resources(plan, flags)


def writeTileSplatFrameUniforms(target, viewProj, plan, splatScale=1.0, shDegree=0, cameraPos=None,
                                viewMatrix=None, projMatrix=None, normalMatrix=None) -> np.ndarray:
    t = write_frame_uniforms(viewProj, plan, splatScale, shDegree, cameraPos, viewMatrix, projMatrix, normalMatrix)
    if target is not None:
        target[:] = t
        return target
    return t


class BindGroups:
    def __init__(self, resources: Resources, keep):
        self.resources = resources
        self._keep = keep


def _host_or_device(arr, dtype):
    """(pointer, memory kind, keepalive) for a numpy array (host) or a torch CUDA tensor (device)."""
    if arr is None:
        return None, None, None
    if hasattr(arr, "data_ptr") and getattr(arr, "is_cuda", False):
        return arr.data_ptr(), _lib.HDSR_MEM_DEVICE, arr
    a = np.ascontiguousarray(arr, dtype=dtype)
    return a.ctypes.data, _lib.HDSR_MEM_HOST, a


def _numel(arr) -> int:
    return int(arr.numel()) if hasattr(arr, "numel") else int(np.asarray(arr).size)


def createTileSplatBindGroups(device, resources: Resources, splatBuffers: dict, outputTexture=None,
                              gbufferTextures=None, *, shDegree: int | None = None) -> BindGroups:
    """splatBuffers: positionBuffer, scaleBuffer, rotationBuffer, opacityBuffer, materialBuffer,
    normalBuffer (optional), shDataBuffer, sortedIndexBuffer; the keys of the reference's call
    (src/splatRenderer.ts:946).  The SH degree is not part of that call: the reference reads it
    from the frame uniforms.  Here it is taken from the shDataBuffer length (N * (6 + 3K) floats)
    unless shDegree is given (keyword or splatBuffers key); the frame uniforms' shDegree must then
    agree.  Every buffer's length is checked against splatCount (ValueError on a mismatch)."""
    n = int(resources.plan["splatCount"])
    deg = shDegree if shDegree is not None else splatBuffers.get("shDegree")
    deg = _lib.HDSR_SH_DEGREE_AUTO if deg is None else int(deg)
    keys = [("positionBuffer", np.float32, 3), ("scaleBuffer", np.float32, 3), ("rotationBuffer", np.float32, 4),
            ("opacityBuffer", np.float32, 1), ("materialBuffer", np.uint32, 1), ("normalBuffer", np.float32, None),
            ("shDataBuffer", np.float32, None), ("sortedIndexBuffer", np.uint32, 1)]
    ptrs, kinds, keep = {}, set(), []
    for k, dt, per in keys:
        arr = splatBuffers.get(k)
        if arr is not None and per is not None and _numel(arr) != per * n:
            raise ValueError(f"{k} holds {_numel(arr)} elements; splatCount {n} needs {per * n}")
        p, kind, ka = _host_or_device(arr, dt)
        ptrs[k] = p
        if kind is not None:
            kinds.add(kind)
        keep.append(ka)
    if len(kinds) > 1:
        raise ValueError("splat buffers must all be host arrays or all device tensors")
    mem = kinds.pop() if kinds else _lib.HDSR_MEM_HOST
    nb = splatBuffers.get("normalBuffer")
    nlen = 0 if nb is None else _numel(nb)  # any length: normal reads are bounded by it
    sb = splatBuffers.get("shDataBuffer")
    s = _lib.Splats(ptrs["positionBuffer"], ptrs["scaleBuffer"], ptrs["rotationBuffer"], ptrs["opacityBuffer"],
                    ptrs["materialBuffer"], ptrs["normalBuffer"], nlen, ptrs["shDataBuffer"],
                    ptrs["sortedIndexBuffer"], n, deg, mem, 0 if sb is None else _numel(sb))
    check(resources._lib.hdsr_bind_splats(resources.handle, C.byref(s)), resources.handle)
    if mem == _lib.HDSR_MEM_DEVICE:
        resources._keep = keep
    return BindGroups(resources, keep)


def encodeFullComputeCompositorPipeline(encoder, resources: Resources, bindGroups: BindGroups | None = None) -> None:
    h = _stream_handle(encoder)
    check(resources._lib.hdsr_render(resources.handle, C.c_void_p(h) if h else None), resources.handle)


def encodeViewDepthSort(encoder, resources: Resources) -> None:
    """The GPU sortedIndices producer (createViewDepthSortInput + encodeGpuSortPrototype,
    src/gpuSortPrototype.ts:120-140, 275-287) for the current frame's view matrix."""
    h = _stream_handle(encoder)
    check(resources._lib.hdsr_sort_back_to_front(resources.handle, C.c_void_p(h) if h else None), resources.handle)


def encodeCompositeOnly(encoder, resources: Resources, bindGroups: BindGroups | None = None) -> None:
    h = _stream_handle(encoder)
    check(resources._lib.hdsr_composite_only(resources.handle, C.c_void_p(h) if h else None), resources.handle)


def bind_scene(resources: Resources, scene, sorted_indices) -> BindGroups:
    """Convenience: binds a hdsr.scenes.Scene (host arrays)."""
    return createTileSplatBindGroups(None, resources, {
        "positionBuffer": scene.positions, "scaleBuffer": scene.scales, "rotationBuffer": scene.rotations,
        "opacityBuffer": scene.opacities, "materialBuffer": scene.materials, "normalBuffer": scene.normals,
        "shDataBuffer": scene.sh_data(), "sortedIndexBuffer": sorted_indices,
    })
