"""Deferred chunk sort, the path of every frame with no debug readback, no G-buffer and no
backward pass: the small tiles and the first 4096-entry chunk of each large tile are sorted, the
rasterizer stops a large tile at the end of that chunk, and only the tiles left unsaturated there
get their other chunks sorted and are resumed from their saved pixel state.  The image must equal
the one-pass render's (a debug context sorts every chunk) bit for bit, and the oracle's; so must
encodeCompositeOnly afterwards.  Two opacity levels: tiles that saturate inside their first chunk,
and tiles that never saturate (every large tile resumed).  Two ways to the first chunks: selected
from each large tile's bucket counts without the bucket scatter (the default after a frame that
resumed less than half of its large-tile references, so the faint scene's later frames on the
same context switch to the other way), and the whole scatter (HDSR_BUCKET_SCATTER)."""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene, frame_inputs
from test_gpu_parity import render_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scatter", [False, True])
@pytest.mark.parametrize("faint", [False, True])
@pytest.mark.parametrize("f16", [False, True])
def test_deferred_equals_one_pass_and_oracle(hip_device, faint, f16, scatter):
    s = dense_scene(n=40_000, seed=17, sh_degree=3, core_fraction=0.6)
    if faint:
        s.opacities *= np.float32(0.05)
    cam = camera.Camera(camera.orbit_eye(2.4, 0.4, 0.25), (0, 0, 0), 420.0, 420.0, 384, 256)
    plan, srt, u = frame_inputs(s, cam)
    full = render_gpu(s, srt, u, plan, flags_debug=True, f16=f16)
    lazy = render_gpu(s, srt, u, plan, flags_debug=False, f16=f16, bucket_scatter=scatter)
    sf, sl = full.stats(), lazy.stats()
    assert sf["chunks"] > sf["large_tiles"] > 0  # large tiles with more than one chunk
    if faint:
        assert sf["composite_max_refs"] > 4096  # tiles resumed in the second round
    rgba, depth = lazy.read_rgba(), lazy.read_depth()
    np.testing.assert_array_equal(rgba, full.read_rgba())
    np.testing.assert_array_equal(depth, full.read_depth())
    assert sl["composite_refs"] == sf["composite_refs"]
    assert sl["composite_max_refs"] == sf["composite_max_refs"]
    if not f16:
        orc = oracle.run(s, srt, u, cam.width, cam.height)
        assert np.abs(rgba.reshape(-1) - orc["rgba"]).max() <= 1e-5
        assert np.abs(depth.reshape(-1) - orc["out_depth"]).max() <= 1e-5
        # the rounds load what one pass loads: whole 256-entry batches up to each tile's saturation
        assert sl["composite_refs"] == int(orc["tile_loaded"].sum())
        assert sl["composite_max_refs"] == int(orc["tile_loaded"].max())
        if faint:  # resumed tiles walk past the entries whose quadrant masks the second round precomputes
            assert int(orc["tile_loaded"].max()) > 4 * 4096
    compositor.encodeCompositeOnly(None, lazy)
    lazy.sync()
    np.testing.assert_array_equal(lazy.read_rgba(), rgba)
    np.testing.assert_array_equal(lazy.read_depth(), depth)
    # reading the entry lists back sorts the chunks the frame skipped: whole lists, as one pass
    np.testing.assert_array_equal(lazy.debug("entries"), full.debug("entries"))
    compositor.encodeCompositeOnly(None, lazy)  # now in one pass over the whole lists
    lazy.sync()
    np.testing.assert_array_equal(lazy.read_rgba(), rgba)
    # further frames on the context (its per-tile flags and saved state are rewritten each frame):
    # the same view again, then another view
    compositor.encodeFullComputeCompositorPipeline(None, lazy)
    lazy.sync()
    np.testing.assert_array_equal(lazy.read_rgba(), rgba)
    assert lazy.stats()["composite_max_refs"] == sf["composite_max_refs"]
    cam2 = camera.Camera(camera.orbit_eye(2.3, 1.1, 0.1), (0, 0, 0), 420.0, 420.0, 384, 256)
    srt2 = scenes.sort_back_to_front(s.positions, cam2.view)
    u2 = camera.frame_uniforms_for(cam2, plan, s.sh_degree)
    full.update_sorted_indices(srt2)
    lazy.update_sorted_indices(srt2)
    for r in (full, lazy):
        r.set_frame(u2)
        compositor.encodeFullComputeCompositorPipeline(None, r)
        r.sync()
    np.testing.assert_array_equal(lazy.read_rgba(), full.read_rgba())
    np.testing.assert_array_equal(lazy.read_depth(), full.read_depth())
    compositor.encodeCompositeOnly(None, lazy)
    lazy.sync()
    np.testing.assert_array_equal(lazy.read_rgba(), full.read_rgba())
    full.destroy()
    lazy.destroy()
