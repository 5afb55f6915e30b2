"""Edge cases of the compute path on the GPU against the oracle: a viewport that is not a multiple
of the tile size, a scene entirely behind the camera, an empty scene, degenerate but finite
attributes (zero-length quaternions, opacity 0 and 1, vanishing and very large scales), and
non-finite attributes (NaN / infinity).  The last are checked for a finite image and the finite
splats' records only: IEEE 754 leaves the sign and payload of a NaN that an operation produces to
the implementation, and the GPU's differ from the host's, so records and per-tile frames that a NaN
reaches are not comparable bit for bit."""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene, frame_inputs
from test_gpu_parity import check_deferred, compare_all, render_gpu

pytestmark = pytest.mark.gpu

BG = np.float32([0.02, 0.02, 0.04, 0.0])


def test_ragged_viewport(hip_device):
    s = dense_scene(n=15_000, seed=21, sh_degree=3, core_fraction=0.3)
    cam = camera.Camera(camera.orbit_eye(2.6, 0.9, 0.35), (0, 0, 0), 310.0, 310.0, 333, 187)
    plan, srt, u = frame_inputs(s, cam)
    assert plan["tileColumns"] * 16 > 333 and plan["tileRows"] * 16 > 187
    orc = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_GBUFFER)
    res = render_gpu(s, srt, u, plan, gbuffer=True)
    compare_all(res, orc, plan, gbuffer=True, exact=True)
    check_deferred(s, srt, u, plan, res, orc)
    res.destroy()


def test_scene_behind_the_camera(hip_device):
    s = scenes.scene_config0(3000, seed=5)
    cam = camera.Camera((0.0, 0.0, 3.0), (0, 0, 0), 400.0, 400.0, 160, 96)
    s.positions[:, 2] += 6.0  # every centre at z >= ~4.5, behind the eye at z = 3 looking down -z
    plan, srt, u = frame_inputs(s, cam)
    orc = oracle.run(s, srt, u, cam.width, cam.height)
    assert orc.stats["visible"] == 0 and orc.stats["refs"] == 0
    res = render_gpu(s, srt, u, plan)
    compare_all(res, orc, plan, exact=True)
    np.testing.assert_array_equal(res.read_rgba().reshape(-1, 4), np.broadcast_to(BG, (160 * 96, 4)))
    check_deferred(s, srt, u, plan, res, orc)
    res.destroy()


def test_empty_scene(hip_device):
    cam = camera.Camera((0.0, 0.0, 3.0), (0, 0, 0), 400.0, 400.0, 100, 70)
    plan = compositor.planTileSplatCompositor({"viewportWidth": 100, "viewportHeight": 70, "tileSizePx": 16,
                                               "splatCount": 0})
    s = scenes.scene_config0(1, seed=1).subset(np.arange(0))
    u = camera.frame_uniforms_for(cam, plan, 0)
    res = render_gpu(s, np.zeros(0, np.uint32), u, plan, flags_debug=False)
    st = res.stats()
    assert st["refs"] == 0 and st["visible_splats"] == 0
    np.testing.assert_array_equal(res.read_rgba().reshape(-1, 4), np.broadcast_to(BG, (100 * 70, 4)))
    np.testing.assert_array_equal(res.read_depth(), np.zeros((70, 100), np.float32))
    res.destroy()


def test_degenerate_finite_attributes(hip_device):
    s = dense_scene(n=8000, seed=33, sh_degree=3, core_fraction=0.4)
    rng = np.random.default_rng(33)
    idx = rng.permutation(s.count)[:1200].reshape(6, 200)
    s.rotations[idx[0]] = 0.0                         # zero-length quaternion (normalised by max(len, 1e-6))
    s.opacities[idx[1]] = 0.0
    s.opacities[idx[2]] = 1.0
    s.scales[idx[3]] = -30.0                          # exp(-30): vanishing footprints
    s.scales[idx[4]] = np.float32([2.0, -6.0, -6.0])  # needles several tiles long
    s.scales[idx[5], 0] = 1.5                         # large anisotropic discs
    cam = camera.Camera(camera.orbit_eye(2.8, 0.3, 0.2), (0, 0, 0), 360.0, 360.0, 320, 240)
    plan, srt, u = frame_inputs(s, cam)
    orc = oracle.run(s, srt, u, cam.width, cam.height)
    res = render_gpu(s, srt, u, plan)
    compare_all(res, orc, plan, exact=True)
    check_deferred(s, srt, u, plan, res, orc)
    res.destroy()


def test_non_finite_attributes(hip_device):
    s = scenes.scene_config0(4000, seed=8)
    bad = np.arange(40, 4000, 97)
    s.positions[bad[0::4]] = np.nan
    s.positions[bad[1::4], 0] = np.inf
    s.scales[bad[2::4]] = np.nan
    s.opacities[bad[3::4]] = np.nan
    cam = camera.Camera(camera.orbit_eye(3.0, 0.4, 0.2), (0, 0, 0), 300.0, 300.0, 256, 160)
    plan = compositor.planTileSplatCompositor({"viewportWidth": 256, "viewportHeight": 160, "tileSizePx": 16,
                                               "splatCount": s.count})
    srt = np.arange(s.count, dtype=np.uint32)  # the host depth sort rejects non-finite depths
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    orc = oracle.run(s, srt, u, cam.width, cam.height)
    res = render_gpu(s, srt, u, plan)
    rgba = res.read_rgba().reshape(-1, 4)
    assert np.all(np.isfinite(rgba)) and np.all(np.isfinite(res.read_depth()))
    assert rgba[:, 3].min() >= 0 and rgba[:, 3].max() <= 1
    finite = np.ones(s.count, bool)
    finite[bad] = False
    np.testing.assert_array_equal(res.debug("projcache").reshape(-1, 14)[finite],
                                  orc["cache"].reshape(-1, 14)[finite])
    check_deferred(s, srt, u, plan, res)
    res.destroy()


def test_tied_keys_and_near_depth_order(hip_device):
    """Tile lists with equal keys (every splat three times at one position) under the depth order,
    and under the depth order with 1% of neighbours swapped: per-tile entries and keys (ties by
    list position) equal the oracle's."""
    base = dense_scene(n=6000, seed=41, sh_degree=0, core_fraction=0.3)
    s = base.subset(np.repeat(np.arange(base.count), 3))
    cam = camera.Camera(camera.orbit_eye(2.7, 0.5, 0.3), (0, 0, 0), 380.0, 380.0, 320, 224)
    plan, srt, u = frame_inputs(s, cam)
    swapped = srt.copy()
    rng = np.random.default_rng(41)
    for i in rng.choice(s.count - 1, s.count // 100, replace=False):
        swapped[i], swapped[i + 1] = swapped[i + 1], swapped[i]
    for order in (srt, swapped):
        orc = oracle.run(s, order, u, cam.width, cam.height)
        res = render_gpu(s, order, u, plan)
        compare_all(res, orc, plan, exact=True)
        check_deferred(s, order, u, plan, res, orc)
        res.destroy()


def test_sh_degree_from_shdata_and_uniform_mismatch(hip_device):
    """The bind takes the SH degree from the shData length when none is given (the reference's bind
    names none, src/splatRenderer.ts:946); frame uniforms with another shDegree are rejected, a
    degree-less bind whose shData matches no degree is rejected, and an explicit degree whose rows
    the shData cannot hold is rejected."""
    from hdsr._lib import HdsrError
    s = dense_scene(n=3000, seed=44, sh_degree=3, core_fraction=0.3)
    cam = camera.Camera(camera.orbit_eye(2.8, 0.3, 0.2), (0, 0, 0), 360.0, 360.0, 192, 128)
    plan, srt, u = frame_inputs(s, cam)
    res = compositor.createTileSplatCompositor(None, plan)
    compositor.bind_scene(res, s, srt)  # no degree given
    bad = u.copy()
    bad.view(np.uint32)[25] = 0
    with pytest.raises(HdsrError, match="differs from the bound shData"):
        res.set_frame(bad)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    res.sync()
    orc = oracle.run(s, srt, u, cam.width, cam.height)
    assert np.abs(res.read_rgba().reshape(-1) - orc["rgba"]).max() <= 1e-5
    bufs = {"positionBuffer": s.positions, "scaleBuffer": s.scales, "rotationBuffer": s.rotations,
            "opacityBuffer": s.opacities, "materialBuffer": s.materials, "sortedIndexBuffer": srt}
    with pytest.raises(HdsrError, match="matches no SH degree"):
        compositor.createTileSplatBindGroups(None, res, {**bufs, "shDataBuffer": s.sh_data()[:-7]})
    with pytest.raises(HdsrError, match="shorter than"):
        compositor.createTileSplatBindGroups(None, res, {**bufs, "shDataBuffer": s.sh_data()[:30 * s.count]},
                                              shDegree=3)
    res.destroy()
