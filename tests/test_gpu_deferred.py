"""The first deferred consumers of the G-buffer on the GPU (hdsr_screen_space_normals,
hdsr_bilateral_normal_filter) against the oracle's restatement, word for word, over the real
first-smoke subset rendered with the G-buffer, in the three normal-recovery modes."""
import numpy as np
import pytest

from hdsr import _lib, assets, camera, compositor, scenes
from oracle import oracle
from smoke_cases import SMOKE, smoke_cameras

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_deferred_normals_match_the_oracle(hip_device, mode):
    a = assets.loadFirstSmokeSplatPayload(SMOKE)
    s = assets.scene_from_attributes(a)
    cam = smoke_cameras(a, width=640, height=480)[1]
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": s.count})
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    res = compositor.createTileSplatCompositor(None, plan, gbuffer=True)
    compositor.bind_scene(res, s, srt)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    H, W = cam.height, cam.width
    depth = res.read_depth()
    src = res.read_output(_lib.OUT_GB_NORMAL, np.uint32, (H, W))
    params = res.normal_params(recovery_mode=mode)
    res.encodeScreenSpaceNormals(None, params)
    ssn = res.read_output(_lib.OUT_FILTERED_NORMAL, np.uint32, (H, W))
    inv, ok = oracle.mat4_inverse(u[:16])
    assert ok
    want = oracle.screen_space_normals(depth, src, inv, u[28:31], mode=mode)
    np.testing.assert_array_equal(ssn, want)
    res.encodeBilateralNormalFilter(None, params)
    got = res.read_output(_lib.OUT_GB_NORMAL, np.uint32, (H, W))
    np.testing.assert_array_equal(got, oracle.bilateral_normals(want, depth))
    covered = depth < 0.9999
    assert covered.mean() > 0.1 and np.any(got[covered] != src[covered])
    res.destroy()
