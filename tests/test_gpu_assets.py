"""The real first-smoke scene (the committed subset of the reference's Scaniverse capture, with
its SAM3 roughness / metalness) decoded from its manifest and sidecars and rendered through the
compositor: every list bit-exact and pixels within 1e-5 of the oracle, G-buffer words included."""

import numpy as np
import pytest

from hdsr import assets, camera, compositor, scenes
from oracle import oracle
from smoke_cases import SMOKE, smoke_cameras

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("view", [0, 1, 2])
def test_smoke_subset_renders_like_the_oracle(hip_device, view):
    a = assets.loadFirstSmokeSplatPayload(SMOKE)
    s = assets.scene_from_attributes(a)
    cam = smoke_cameras(a)[view]
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": s.count})
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    orc = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_GBUFFER)
    res = compositor.createTileSplatCompositor(None, plan, debug=True, gbuffer=True)
    compositor.bind_scene(res, s, srt)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    res.sync()
    from test_gpu_parity import compare_all
    compare_all(res, orc, plan, gbuffer=True)
    assert orc.stats["visible"] > 0.5 * s.count and orc.stats["refs"] > s.count
    print(f"smoke view {view}: refs={orc.stats['refs']} large_tiles={orc.stats['large_tiles']}")
    res.destroy()
