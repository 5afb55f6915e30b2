"""Small seeded scenes for parity tests (sizes the oracle finishes in seconds)."""
from __future__ import annotations

import numpy as np

from hdsr import camera, compositor, scenes


def dense_scene(n=40_000, seed=7, sh_degree=0, core_fraction=0.6, core_sigma=0.03):
    """A config-0 style scene with a dense core so several tiles exceed 4096 entries."""
    s = scenes.scene_config0(n, seed)
    rng = np.random.default_rng(seed + 100)
    k = int(n * core_fraction)
    s.positions[:k] = rng.normal(0.0, core_sigma, (k, 3)).astype(np.float32)
    if sh_degree:
        kk = (sh_degree + 1) ** 2 - 1
        s.sh = rng.normal(0.0, 0.05, (n, kk, 3)).astype(np.float32)
        s.sh_degree = sh_degree
    return s


def frame_inputs(scene, cam, sorted_mode="depth"):
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": scene.count})
    if sorted_mode == "depth":
        srt = scenes.sort_back_to_front(scene.positions, cam.view)
    else:
        srt = np.arange(scene.count, dtype=np.uint32)
    u = camera.frame_uniforms_for(cam, plan, scene.sh_degree)
    return plan, srt, u
