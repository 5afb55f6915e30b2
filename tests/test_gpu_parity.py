"""GPU parity: every intermediate and output of the HIP path against the CPU oracle.

Integer outputs (tile boxes, tile counts / offsets, canonical and sorted entry lists, 20-bit keys,
tile lists, chunks) must be bit-exact.  The projection records are bit-exact too (both sides use
IEEE fp32 without contraction and the same exp/ln).  Pixels: max |GPU - oracle| <= 1e-5 in linear
RGBA (the build reaches 0), and the error of both against the FP64 twin is reported.
"""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene, frame_inputs

pytestmark = pytest.mark.gpu

PIXEL_TOL = 1e-5


def render_gpu(scene, srt, u, plan, flags_debug=True, gbuffer=False, f16=False, bucket_scatter=False):
    res = compositor.createTileSplatCompositor(None, plan, {"f16": f16}, debug=flags_debug, gbuffer=gbuffer,
                                               bucket_scatter=bucket_scatter)
    compositor.bind_scene(res, scene, srt)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    res.sync()
    return res


def compare_all(res, orc, plan, gbuffer=False, exact=False):
    st = res.stats()
    assert st["refs"] == orc.stats["refs"]
    assert st["large_splats"] == orc.stats["large_splats"]
    assert st["visible_splats"] == orc.stats["visible"]
    assert st["small_tiles"] == orc.stats["small_tiles"]
    assert st["large_tiles"] == orc.stats["large_tiles"]
    assert st["chunks"] == orc.stats["chunks"]
    cache = res.debug("projcache")
    ocache = orc["cache"]
    mism = np.nonzero(cache != ocache)[0]
    assert mism.size == 0, f"projcache words differ at {mism[:10]} (rank {mism[:10] // 14}, word {mism[:10] % 14})"
    np.testing.assert_array_equal(res.debug("depth"), orc["depth"])
    np.testing.assert_array_equal(res.debug("large_splats"), orc["large_list"])
    np.testing.assert_array_equal(res.debug("tile_counts"), orc["tile_counts"])
    np.testing.assert_array_equal(res.debug("tile_offsets"), orc["tile_offsets"])
    np.testing.assert_array_equal(res.debug("entries_canon"), orc["entries_canon"])
    np.testing.assert_array_equal(res.debug("small_list"), orc["small_tiles"])
    np.testing.assert_array_equal(res.debug("large_list"), orc["large_tiles"])
    np.testing.assert_array_equal(res.debug("large_bases"), orc["large_bases"])
    np.testing.assert_array_equal(res.debug("chunks"), orc["chunks"])
    np.testing.assert_array_equal(res.debug("keys20"), orc["keys"])
    np.testing.assert_array_equal(res.debug("entries"), orc["entries"])
    rgba = res.read_rgba().reshape(-1)
    err = np.abs(rgba - orc["rgba"]).max()
    assert err <= PIXEL_TOL, f"max pixel error {err}"
    depth = res.read_depth().reshape(-1)
    assert np.abs(depth - orc["out_depth"]).max() <= PIXEL_TOL
    if exact:
        np.testing.assert_array_equal(rgba, orc["rgba"])
        np.testing.assert_array_equal(depth, orc["out_depth"])
    if gbuffer:
        H, W = plan["viewportHeight"], plan["viewportWidth"]
        from hdsr import _lib
        gn = res.read_output(_lib.OUT_GB_NORMAL, np.uint32, (H * W,))
        gm = res.read_output(_lib.OUT_GB_MATERIAL, np.uint32, (H * W * 4,))
        np.testing.assert_array_equal(gn, orc["gb_normal"])
        np.testing.assert_array_equal(gm, orc["gb_material"])
    return err


def check_deferred(scene, srt, u, plan, one_pass, orc=None, f16=False):
    """Renders the frame on contexts with no debug readback and no G-buffer, which run the
    deferred chunk sort (first chunks, rasterize, sort the resumed tiles' other chunks, resume),
    once selecting each large tile's first chunk without the bucket scatter (a fresh context's
    default) and once with the scatter (HDSR_BUCKET_SCATTER), and requires each image and depth to
    equal the one-pass context's bit for bit (and the oracle's within 1e-5 when given); the entry
    lists each completes on readback equal the one-pass lists."""
    sf = one_pass.stats()
    for scatter in (False, True):
        lazy = render_gpu(scene, srt, u, plan, flags_debug=False, f16=f16, bucket_scatter=scatter)
        rgba, depth = lazy.read_rgba(), lazy.read_depth()
        np.testing.assert_array_equal(rgba, one_pass.read_rgba())
        np.testing.assert_array_equal(depth, one_pass.read_depth())
        if orc is not None:
            assert np.abs(rgba.reshape(-1) - orc["rgba"]).max() <= PIXEL_TOL
            assert np.abs(depth.reshape(-1) - orc["out_depth"]).max() <= PIXEL_TOL
        sl = lazy.stats()
        assert sl["refs"] == sf["refs"] and sl["composite_refs"] == sf["composite_refs"]
        np.testing.assert_array_equal(lazy.debug("entries"), one_pass.debug("entries"))
        lazy.destroy()
    return sl


def test_config0_full_path(hip_device):
    s = scenes.scene_config0()
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    orc = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_FP64)
    res = render_gpu(s, srt, u, plan)
    err = compare_all(res, orc, plan)
    fp64_err = np.abs(res.read_rgba().reshape(-1) - orc["rgba64"]).max()
    print(f"config0: max|gpu-oracle|={err:.3g}, max|gpu-fp64|={fp64_err:.3g}, refs={orc.stats['refs']}")
    res.destroy()


def test_dense_scene_bucket_cascade(hip_device):
    s = dense_scene()
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    orc = oracle.run(s, srt, u, cam.width, cam.height)
    assert orc.stats["large_tiles"] > 0 and orc.stats["chunks"] > orc.stats["large_tiles"]
    res = render_gpu(s, srt, u, plan)
    compare_all(res, orc, plan)
    res.destroy()


def test_sh3_gbuffer_identity_order(hip_device):
    s = dense_scene(n=20_000, seed=11, sh_degree=3, core_fraction=0.3)
    cam = camera.Camera(camera.orbit_eye(2.5, 0.7, 0.3), (0, 0, 0), 700.0, 700.0, 600, 420)
    plan, srt, u = frame_inputs(s, cam, sorted_mode="identity")
    orc = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_GBUFFER)
    res = render_gpu(s, srt, u, plan, gbuffer=True)
    compare_all(res, orc, plan, gbuffer=True)
    res.destroy()


def test_repeat_runs_bit_identical(hip_device):
    s = dense_scene(n=30_000, seed=3)
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    res = render_gpu(s, srt, u, plan)
    a = (res.debug("entries"), res.debug("keys20"), res.read_rgba())
    for _ in range(3):
        compositor.encodeFullComputeCompositorPipeline(None, res)
        res.sync()
        b = (res.debug("entries"), res.debug("keys20"), res.read_rgba())
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    res.destroy()


@pytest.mark.parametrize("gbuffer", [False, True])
def test_f16_composite_bit_exact(hip_device, gbuffer):
    """options.f16 (gpu_tile_splat_composite_f16.wgsl): native binary16 transmittance / colour on
    the GPU against the oracle's exactly-rounded restatement, bit for bit (lists, image, depth,
    G-buffer words)."""
    s = dense_scene(n=20_000, seed=12, sh_degree=3, core_fraction=0.5)
    cam = camera.Camera(camera.orbit_eye(2.5, 0.6, 0.2), (0, 0, 0), 700.0, 700.0, 600, 420)
    plan, srt, u = frame_inputs(s, cam)
    flags = oracle.FLAG_F16 | oracle.FLAG_FP64 | (oracle.FLAG_GBUFFER if gbuffer else 0)
    orc = oracle.run(s, srt, u, cam.width, cam.height, flags)
    res = render_gpu(s, srt, u, plan, gbuffer=gbuffer, f16=True)
    compare_all(res, orc, plan, gbuffer=gbuffer, exact=True)
    fp64_err = np.abs(res.read_rgba().reshape(-1) - orc["rgba64"]).max()
    print(f"f16 composite: bit-exact vs oracle, max|gpu-fp64|={fp64_err:.3g}")
    res.destroy()
