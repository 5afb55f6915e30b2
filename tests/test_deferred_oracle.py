"""The oracle's deferred consumers of the G-buffer (screen-space normals, bilateral normal filter,
mat4Inverse), pinned by what geometry fixes: a plane's depth image gives back the plane's normal,
a constant normal field passes the filter unchanged, normals are not averaged across a depth step,
and the inverse matches numpy's."""
import math

import numpy as np

from hdsr import camera
from oracle import oracle


def test_mat4_inverse_matches_numpy_and_rejects_singular():
    rng = np.random.default_rng(2)
    for _ in range(20):
        m = rng.normal(0, 1, (4, 4)).astype(np.float32)
        inv, ok = oracle.mat4_inverse(m.T.reshape(16))  # column-major storage
        assert ok
        want = np.linalg.inv(m.astype(np.float64)).T.reshape(16)
        np.testing.assert_allclose(inv, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())
    inv, ok = oracle.mat4_inverse(np.zeros(16, np.float32))
    assert not ok


def plane_depth(vp16, n, p0, W, H):
    """NDC depth of the plane n.x = n.p0 along every pixel-centre ray (float64)."""
    vp = camera.col_major(vp16).astype(np.float64)
    inv = np.linalg.inv(vp)
    xs, ys = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    ndc = np.stack([xs / W * 2 - 1, 1 - ys / H * 2], -1)
    a = inv @ np.concatenate([ndc, np.zeros((H, W, 1)), np.ones((H, W, 1))], -1)[..., None]
    b = inv @ np.concatenate([ndc, np.ones((H, W, 1)), np.ones((H, W, 1))], -1)[..., None]
    a = a[..., :3, 0] / a[..., 3:, 0]
    b = b[..., :3, 0] / b[..., 3:, 0]
    t = (np.dot(p0, n) - a @ n) / ((b - a) @ n)
    x = a + t[..., None] * (b - a)
    c = vp @ np.concatenate([x, np.ones((H, W, 1))], -1)[..., None]
    return (c[..., 2, 0] / c[..., 3, 0]).astype(np.float32)


def test_screen_space_normals_of_a_plane():
    # near 1 / far 50 keeps the fp32 reconstruction well conditioned: with the viewer's 0.01 / 100
    # the depths of a surface 3 units away sit within 0.3% of 1 and the shader's fp32 w term
    # cancels (the reference's own behaviour, reproduced exactly by the restatement)
    W, H = 96, 64
    cam = camera.Camera((0.3, 0.5, 3.0), (0, 0, 0), 80.0, 80.0, W, H)
    vp16 = camera.view_projection(camera.pinhole_projection(80.0, 80.0, W, H, near=1.0, far=50.0), cam.view)
    n = np.array([0.2, 0.9, 0.35])
    n /= np.linalg.norm(n)
    depth = plane_depth(vp16, n, np.zeros(3), W, H)
    inv, ok = oracle.mat4_inverse(vp16)
    assert ok
    eye = np.asarray(cam.eye, np.float32)
    src = np.zeros((H, W), np.uint32)
    out = oracle.screen_space_normals(depth, src, inv, eye, mode=0)
    got = oracle.oct_words_to_normals(out)
    facing = n if np.dot(n, eye) > 0 else -n  # the plane passes through the origin
    # interior pixels (on the image border the clamped neighbour repeats the centre's depth and wins
    # the smaller-difference choice: the reference's own edge behaviour)
    ang = np.degrees(np.arccos(np.clip(got[1:-1, 1:-1] @ facing, -1, 1)))
    assert np.median(ang) < 0.1 and ang.max() < 1.0, (np.median(ang), ang.max())
    # pass-through mode: the source normal turned towards the camera
    src = np.full((H, W), oracle.normals_to_oct_words(-facing), np.uint32)
    out = oracle.screen_space_normals(depth, src, inv, eye, mode=2)
    got = oracle.oct_words_to_normals(out)
    assert np.degrees(np.arccos(np.clip(got @ facing, -1, 1))).max() < 0.2
    # recovery mode: grazing bounds that force the smoothstep to 1 keep the source normal, bounds that
    # force it to 0 give the screen-space normal (a swapped blend fails one of the two)
    keep = oracle.screen_space_normals(depth, src, inv, eye, mode=1, grazing_start=-2.0, grazing_end=-1.0)
    np.testing.assert_array_equal(keep, oracle.screen_space_normals(depth, src, inv, eye, mode=2))
    rec = oracle.oct_words_to_normals(oracle.screen_space_normals(depth, src, inv, eye, mode=1,
                                                                 grazing_start=2.0, grazing_end=3.0))
    ang = np.degrees(np.arccos(np.clip(rec[1:-1, 1:-1] @ facing, -1, 1)))
    assert np.median(ang) < 0.1 and ang.max() < 1.0
    # background pixels: a zero word in screen-space mode, the source word in recovery modes
    depth[0, 0] = 1.0
    assert oracle.screen_space_normals(depth, src, inv, eye, mode=0)[0, 0] == 0
    assert oracle.screen_space_normals(depth, src, inv, eye, mode=1)[0, 0] == src[0, 0]


def test_bilateral_filter_keeps_constant_fields_and_depth_edges():
    H, W = 40, 48
    a = np.array([0.0, 0.6, 0.8])
    b = np.array([0.8, 0.0, 0.6])
    words = np.full((H, W), oracle.normals_to_oct_words(a), np.uint32)
    depth = np.full((H, W), 0.5, np.float32) + np.linspace(0, 1e-3, W, dtype=np.float32)[None, :]
    out = oracle.bilateral_normals(words, depth)
    np.testing.assert_allclose(oracle.oct_words_to_normals(out), np.broadcast_to(a, (H, W, 3)), atol=2e-3)
    # two regions with different normals, separated by a 2x depth step: no blending across it
    words[:, W // 2:] = oracle.normals_to_oct_words(b)
    depth[:, W // 2:] = 0.9
    out = oracle.oct_words_to_normals(oracle.bilateral_normals(words, depth))
    assert np.degrees(np.arccos(np.clip(out[:, W // 2 - 1] @ a, -1, 1))).max() < 0.5
    assert np.degrees(np.arccos(np.clip(out[:, W // 2] @ b, -1, 1))).max() < 0.5
    # the same step in normals at equal depth is smoothed (the depth weight is what keeps edges)
    depth[:, W // 2:] = depth[:, :W // 2].max()
    out = oracle.oct_words_to_normals(oracle.bilateral_normals(words, depth, normal_sigma=100.0))
    assert np.degrees(np.arccos(np.clip(out[:, W // 2 - 1] @ a, -1, 1))).min() > 5.0
    # background passes through
    depth[3, 4] = 1.0
    assert oracle.bilateral_normals(words, depth)[3, 4] == words[3, 4]
