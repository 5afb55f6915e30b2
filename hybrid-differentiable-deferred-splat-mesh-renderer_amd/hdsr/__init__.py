"""hdsr - MI355X (gfx950) implementation of the compute-compositor splat path.

The compute path is libhdsr.so (hand-written HIP kernels behind the C ABI of include/hdsr.h);
this package is its Python host side: a facade with the reference's gpuTileSplatCompositor
names, camera / uniform helpers and seeded synthetic scenes.
"""
from . import camera, scenes
from ._lib import HdsrError, load
from .compositor import (TILE_SPLAT_FRAME_UNIFORM_BYTES, BindGroups, Resources, bind_scene,
                         createTileSplatBindGroups, createTileSplatCompositor, encodeCompositeOnly,
                         encodeFullComputeCompositorPipeline, planTileSplatCompositor,
                         writeTileSplatFrameUniforms)

__all__ = [
    "camera", "scenes", "HdsrError", "load", "TILE_SPLAT_FRAME_UNIFORM_BYTES", "BindGroups", "Resources",
    "bind_scene", "createTileSplatBindGroups", "createTileSplatCompositor", "encodeCompositeOnly",
    "encodeFullComputeCompositorPipeline", "planTileSplatCompositor", "writeTileSplatFrameUniforms",
]
