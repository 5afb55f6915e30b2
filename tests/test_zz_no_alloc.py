"""hdsr_sort_back_to_front_noalloc (the allocation-free depth order: fixed device scratch, caller's
input and output buffers) against hdsr_sort_back_to_front on a context, bit for bit, on two calls
in a row with the second smaller than the first (the scratch keeps the first call's contents),
and the oracle's restatement of the reference's network; sizes beyond the scratch are rejected
before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from hdsr import _lib, camera, compositor, scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def order_via_context(sc, cam):
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": sc.count})
    res = compositor.createTileSplatCompositor(None, plan)
    compositor.bind_scene(res, sc, np.arange(sc.count, dtype=np.uint32))
    u = camera.frame_uniforms_for(cam, plan, sc.sh_degree)
    res.set_frame(u)
    compositor.encodeViewDepthSort(None, res)
    out = res.debug("sorted")
    res.destroy()
    return out, u


def order_noalloc(sc, u):
    import torch
    lib = _lib.load()
    pos = torch.from_numpy(np.ascontiguousarray(sc.positions, np.float32)).cuda()
    out = torch.full((sc.count,), 0xABCDEF, dtype=torch.int32, device="cuda")
    uu = np.ascontiguousarray(u, np.float32)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.hdsr_sort_back_to_front_noalloc(pos.data_ptr(), sc.count, uu.ctypes.data, out.data_ptr(), stream)
    assert rc == 0, lib.hdsr_last_error(None)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def test_noalloc_depth_sort_equals_context_sort(hip_device):
    cam = scenes.cameras_config0()[0]
    for n, seed in [(60_000, 1), (3_001, 2)]:  # the second call is smaller: stale scratch must not leak
        sc = scenes.scene_config0(n, seed=seed)
        sc.positions[n // 3: n // 3 + 500] = sc.positions[:500]  # exact key ties
        want, u = order_via_context(sc, cam)
        got = order_noalloc(sc, u)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, orc.gpu_sort_order(orc.view_depth_keys_f32(sc.positions, cam.view)))


def test_noalloc_rejects_oversize_and_null(hip_device):
    import torch
    lib = _lib.load()
    u = np.zeros(64, np.float32)
    buf = torch.zeros(16, dtype=torch.float32, device="cuda")
    rc = lib.hdsr_sort_back_to_front_noalloc(buf.data_ptr(), _lib.HDSR_NOALLOC_MAX_SPLATS + 1, u.ctypes.data,
                                             buf.data_ptr(), None)
    assert rc == _lib.HDSR_E_CAPACITY and b"exceed the scratch" in lib.hdsr_last_error(None)
    assert lib.hdsr_sort_back_to_front_noalloc(None, 5, u.ctypes.data, buf.data_ptr(), None) == _lib.HDSR_E_INVALID_ARG
    assert lib.hdsr_sort_back_to_front_noalloc(None, 0, u.ctypes.data, None, None) == 0
