"""The reference's WebGPU caps on the GPU (HDSR_REF_CAPS / options.refCaps) against the oracle's
same mode (FLAG_REF_CAPS, pinned in test_ref_caps_oracle.py), at sizes that trip each cap:

  * the reference budget maxTotalTileRefs alone (a dense 20k-splat scene, tileEntryMultiplier 2);
  * more than 65,535 large splats with the budget (100k splats of about 12 x 12 tiles, multiplier
    40), and with more than 1,024 chunks besides (multiplier 200: the 65535 x 256 dispatch cap);
  * no cap tripped (config-0 scale): the capped context renders exactly as the lifted one.

Bit-exact: tile counts and offsets, the stored part of every tile's list and keys, the small /
large tile lists, overflow bases and chunks; pixels and depth within 1e-5 (in practice 0).
"""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene

pytestmark = pytest.mark.gpu


def big_scene(n=100_000, seed=5, sigma=0.3):
    s = scenes.scene_config0(n, seed)
    rng = np.random.default_rng(seed)
    s.positions = (rng.uniform(-1, 1, (n, 3)) * np.array([1.2, 0.9, 0.3])).astype(np.float32)
    s.scales = (np.full((n, 3), np.log(sigma)) + rng.normal(0, 0.2, (n, 3))).astype(np.float32)
    return s


CASES = {
    "budget": (lambda: dense_scene(n=20_000, seed=6, core_fraction=0.3), 320, 240, 2.0),
    "large_list": (big_scene, 512, 384, 40.0),
    "all_caps": (big_scene, 512, 384, 200.0),
}


def render(s, srt, u, plan, ref_caps, debug):
    res = compositor.createTileSplatCompositor(None, plan, {"refCaps": ref_caps}, debug=debug)
    compositor.bind_scene(res, s, srt)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    res.sync()
    return res


@pytest.mark.parametrize("case", list(CASES))
def test_ref_caps_match_oracle(hip_device, case):
    make, W, H, mult = CASES[case]
    s = make()
    cam = camera.Camera((0, 0, 3.0), (0, 0, 0), 300.0, 300.0, W, H)
    plan = compositor.planTileSplatCompositor({"viewportWidth": W, "viewportHeight": H, "tileSizePx": 16,
                                               "splatCount": s.count, "tileEntryMultiplier": mult})
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    orc = oracle.run(s, srt, u, W, H, oracle.FLAG_REF_CAPS)
    print(case, orc.stats)
    assert orc.stats["cap_refs_tripped"] == 1
    if case != "budget":
        assert orc.stats["cap_large_splats_tripped"] == 1
    if case == "all_caps":
        assert orc.stats["cap_chunks_tripped"] == 1
    res = render(s, srt, u, plan, True, True)
    st = res.stats()
    for k, ko in (("refs", "refs"), ("large_splats", "large_splats"), ("small_tiles", "small_tiles"),
                  ("large_tiles", "large_tiles"), ("chunks", "chunks")):
        assert st[k] == orc.stats[ko], (k, st[k], orc.stats[ko])
    assert st["overflow_reruns"] == 0
    offs, cnts = res.debug("tile_offsets"), res.debug("tile_counts")
    np.testing.assert_array_equal(offs, orc["tile_offsets"])
    np.testing.assert_array_equal(cnts, orc["tile_counts"])
    ent, keys = res.debug("entries"), res.debug("keys20")
    for m in np.nonzero(cnts)[0]:
        a, c = int(offs[m]), int(cnts[m])
        np.testing.assert_array_equal(ent[a:a + c], orc["entries"][a:a + c], err_msg=f"tile {m}")
        np.testing.assert_array_equal(keys[a:a + c], orc["keys"][a:a + c], err_msg=f"tile {m} keys")
    np.testing.assert_array_equal(res.debug("small_list"), orc["small_tiles"])
    np.testing.assert_array_equal(res.debug("large_list"), orc["large_tiles"])
    np.testing.assert_array_equal(res.debug("large_bases"), orc["large_bases"])
    np.testing.assert_array_equal(res.debug("chunks"), orc["chunks"])
    rgba, depth = res.read_rgba().reshape(-1), res.read_depth().reshape(-1)
    err = np.abs(rgba - orc["rgba"]).max()
    assert err <= 1e-5 and np.abs(depth - orc["out_depth"]).max() <= 1e-5, err
    # the same frame without debug readback renders the same image
    plain = render(s, srt, u, plan, True, False)
    np.testing.assert_array_equal(plain.read_rgba().reshape(-1), rgba)
    res.destroy()
    plain.destroy()


def test_ref_caps_without_trips_equal_lifted(hip_device):
    s = dense_scene(n=20_000, seed=4)
    W, H = 320, 240
    cam = camera.Camera((0, 0, 3.0), (0, 0, 0), 300.0, 300.0, W, H)
    plan = compositor.planTileSplatCompositor({"viewportWidth": W, "viewportHeight": H, "tileSizePx": 16,
                                               "splatCount": s.count})
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    capped = render(s, srt, u, plan, True, True)
    lifted = render(s, srt, u, plan, False, True)
    st = capped.stats()
    assert not (st["cap_refs_tripped"] or st["cap_large_splats_tripped"] or st["cap_chunks_tripped"])
    for k in ("tile_counts", "tile_offsets", "entries", "keys20", "chunks"):
        np.testing.assert_array_equal(capped.debug(k), lifted.debug(k), err_msg=k)
    np.testing.assert_array_equal(capped.read_rgba(), lifted.read_rgba())
    np.testing.assert_array_equal(capped.read_depth(), lifted.read_depth())
    capped.destroy()
    lifted.destroy()
