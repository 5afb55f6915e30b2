"""The GPU sortedIndices producer (hdsr_sort_back_to_front) against the oracle's restatement of
the reference's bitonic network (oracle.gpu_sort_order, pinned in test_host.py against the
network run pair by pair): bit-exact orders, including exact key ties, sizes that are not powers
of two, a single splat, and the 1M-splat view; a frame rendered from the GPU order equals the frame
rendered from the same order supplied by the host."""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def gpu_order(scene, cam):
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": scene.count})
    res = compositor.createTileSplatCompositor(None, plan)
    compositor.bind_scene(res, scene, np.arange(scene.count, dtype=np.uint32))
    res.set_frame(camera.frame_uniforms_for(cam, plan, scene.sh_degree))
    compositor.encodeViewDepthSort(None, res)
    out = res.debug("sorted")
    return res, plan, out


@pytest.mark.parametrize("n,seed", [(1, 1), (3, 2), (4097, 3), (50_000, 4)])
def test_depth_sort_matches_network(hip_device, n, seed):
    sc = scenes.scene_config0(n, seed=seed)
    cam = scenes.cameras_config0()[0]
    res, _, got = gpu_order(sc, cam)
    want = orc.gpu_sort_order(orc.view_depth_keys_f32(sc.positions, cam.view))
    np.testing.assert_array_equal(got, want)
    res.destroy()


def test_depth_sort_exact_ties(hip_device):
    # duplicated positions give exactly equal keys: the order is the network's, not the index order
    sc = scenes.scene_config0(3000, seed=5)
    sc.positions[1000:2000] = sc.positions[:1000]
    sc.positions[2000:2500] = sc.positions[:500] * np.float32(1.0)
    cam = scenes.cameras_config0()[0]
    res, _, got = gpu_order(sc, cam)
    keys = orc.view_depth_keys_f32(sc.positions, cam.view)
    want = orc.gpu_sort_order(keys)
    np.testing.assert_array_equal(got, want)
    assert not np.array_equal(want, np.lexsort((np.arange(len(keys)), -keys.astype(np.float64))))
    res.destroy()


def test_depth_sort_1m_view_and_render(hip_device):
    sc = scenes.scene_config1(1_000_000)
    cam = scenes.cameras_config1()[3]
    res, plan, got = gpu_order(sc, cam)
    want = orc.gpu_sort_order(orc.view_depth_keys_f32(sc.positions, cam.view))
    np.testing.assert_array_equal(got, want)
    # render from the GPU order, then from the same order given by the host
    compositor.encodeFullComputeCompositorPipeline(None, res)
    a = res.read_rgba()
    res.update_sorted_indices(want)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    np.testing.assert_array_equal(res.read_rgba(), a)
    res.destroy()
