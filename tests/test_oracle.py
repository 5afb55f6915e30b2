"""Pins the CPU oracle independently of the kernels.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * structural invariants of binning (offsets = exclusive scan of counts, every visible splat in
    exactly the tiles of its box, canonical small-then-large rank order per tile, row-major tile
    lists, overflow bases as a running sum, chunks = consecutive 4096-entry ranges);
  * the per-tile order: keys strictly increasing, q20 >> 13 equals the bucket computed in float64
    from the depths (up to ulp-boundary cases), stable on equal q20;
  * projection closed forms in float64: conic = inverse of the low-passed 2D covariance, the 3D
    covariance rebuilt from the quaternion, tile box from the 3-sigma radius, SH degree 0 = DC;
  * compositing closed forms: a single isotropic splat at a pixel centre gives
    alpha = min(0.99, o * exp(power)) exactly; a brute-force per-pixel front-to-back blend over
    the oracle's own order equals the oracle image; the FP64 twin stays close.
"""
import math

import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene, frame_inputs


@pytest.fixture(scope="module")
def small_frame():
    s = scenes.scene_config0(n=4000, seed=5)
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    return s, cam, plan, srt, u, oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_FP64)


@pytest.fixture(scope="module")
def dense_frame():
    s = dense_scene(n=12_000, seed=9, core_fraction=0.8, core_sigma=0.02)
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    return s, cam, plan, srt, u, oracle.run(s, srt, u, cam.width, cam.height)


def boxes(cache):
    w7 = cache.reshape(-1, 14)[:, 7]
    vis = w7 != 0xFFFFFFFF
    x0, y0 = w7 & 0xFF, (w7 >> 8) & 0xFF
    x1, y1 = (w7 >> 16) & 0xFF, w7 >> 24
    return vis, x0, y0, x1, y1


def morton(x, y):
    x = np.asarray(x, np.uint64)
    y = np.asarray(y, np.uint64)
    out = np.zeros_like(x)
    for b in range(16):
        out |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        out |= ((y >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    return out


@pytest.mark.parametrize("which", ["small_frame", "dense_frame"])
def test_binning_invariants(which, request):
    s, cam, plan, srt, u, f = request.getfixturevalue(which)
    counts, offs = f["tile_counts"], f["tile_offsets"]
    assert counts.size == plan["mortonTileCount"]
    np.testing.assert_array_equal(offs, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))
    R = int(counts.sum())
    assert R == f.stats["refs"] == f["entries"].size
    vis, x0, y0, x1, y1 = boxes(f["cache"])
    span = (x1.astype(np.int64) - x0 + 1) * (y1.astype(np.int64) - y0 + 1)
    assert int(span[vis].sum()) == R
    large = vis & (span > 16)
    np.testing.assert_array_equal(f["large_list"], np.nonzero(large)[0])
    # expected per-tile canonical list from the boxes: small ranks ascending then large ranks ascending
    tiles = {}
    for r in list(np.nonzero(vis & ~large)[0]) + list(np.nonzero(large)[0]):
        for ty in range(y0[r], y1[r] + 1):
            for tx in range(x0[r], x1[r] + 1):
                tiles.setdefault(int(morton(tx, ty)), []).append(int(r))
    canon = f["entries_canon"]
    for m, lst in tiles.items():
        np.testing.assert_array_equal(canon[offs[m]:offs[m] + counts[m]], lst)
    # final entries are a per-tile permutation of the canonical ones
    for m in np.nonzero(counts)[0]:
        a = np.sort(canon[offs[m]:offs[m] + counts[m]])
        b = np.sort(f["entries"][offs[m]:offs[m] + counts[m]])
        np.testing.assert_array_equal(a, b)
    # tile lists in row-major order, overflow bases as a running sum past R
    rowmajor = [int(morton(i % plan["tileColumns"], i // plan["tileColumns"])) for i in range(plan["tileCount"])]
    small = [m for m in rowmajor if 0 < counts[m] <= 4096]
    large_t = [m for m in rowmajor if counts[m] > 4096]
    np.testing.assert_array_equal(f["small_tiles"], small)
    np.testing.assert_array_equal(f["large_tiles"], large_t)
    bases = R + np.concatenate([[0], np.cumsum([counts[m] for m in large_t])[:-1]]) if large_t else []
    np.testing.assert_array_equal(f["large_bases"], np.asarray(bases, np.uint32))
    ch = f["chunks"].reshape(-1, 2)
    exp = [(offs[m] + c0, min(4096, counts[m] - c0)) for m in large_t for c0 in range(0, counts[m], 4096)]
    np.testing.assert_array_equal(ch, np.asarray(exp, np.uint32).reshape(-1, 2))


def _log_frame64(d32):
    lmin = math.log(max(float(d32.min()), 1e-6))
    lr = math.log(max(float(d32.max()), 1e-6)) - lmin
    return lmin, (0.0 if lr < 1e-10 else 128.0 / lr)


@pytest.mark.parametrize("which", ["small_frame", "dense_frame"])
def test_per_tile_keys_and_order(which, request):
    s, cam, plan, srt, u, f = request.getfixturevalue(which)
    counts, offs, keys, ent = f["tile_counts"], f["tile_offsets"], f["keys"], f["entries"]
    depth = f["depth"].view(np.float32)
    chunk_of = {}
    for start, cnt in f["chunks"].reshape(-1, 2):
        chunk_of[int(start)] = int(cnt)
    near_boundary = 0
    for m in np.nonzero(counts)[0]:
        lo, c = int(offs[m]), int(counts[m])
        d_tile = depth[ent[lo:lo + c]]
        lmin, scale = _log_frame64(d_tile)
        # fp32 ln is within 2 ulp; the key scale 2^13 * 128 / logRange amplifies that error
        lmag = max(abs(lmin), abs(math.log(max(float(d_tile.max()), 1e-6))))
        tol = 1 + math.ceil(8192 * scale * 6 * float(np.spacing(np.float32(lmag))))
        segs = [(lo, c)] if c <= 4096 else [(a, chunk_of[a]) for a in range(lo, lo + c, 4096)]
        for a, n in segs:
            k = keys[a:a + n]
            assert np.all(np.diff(k.astype(np.int64)) > 0), "keys not strictly increasing"
            assert np.all((k & 0xFFF) < n)
            assert len(np.unique(k & 0xFFF)) == n, "local indices not a permutation"
            q20 = k >> 12
            y64 = (np.log(np.maximum(depth[ent[a:a + n]].astype(np.float64), 1e-6)) - lmin) * scale
            q64 = np.clip(np.floor(y64 * 8192), 0, 2**20 - 1)
            diff = np.abs(q64 - q20)
            near_boundary += int((diff >= 1).sum())
            assert diff.max() <= tol, "q20 disagrees with the float64 log-depth position"
        if c > 4096:
            # bucket order across the tile: q20 >> 13 non-decreasing over the chunk sequence
            b = keys[lo:lo + c] >> 25
            firsts = [b[i:i + 4096] for i in range(0, c, 4096)]
            for x, y in zip(firsts, firsts[1:]):
                assert x.min() <= y.min() and x.max() <= y.max()
    # most keys agree exactly with the float64 evaluation (differences come from narrow tiles)
    assert near_boundary <= 0.1 * max(1, f.stats["refs"])


def test_dense_frame_exercises_cascade(dense_frame):
    f = dense_frame[-1]
    assert f.stats["large_tiles"] >= 1
    assert f.stats["chunks"] > f.stats["large_tiles"]


def _cov3(scale_log, q):
    w, x, y, z = (q / np.linalg.norm(q)).astype(np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    S = np.diag(np.exp(scale_log.astype(np.float64)))
    return R @ S @ S @ R.T


def test_projection_closed_forms(small_frame):
    s, cam, plan, srt, u, f = small_frame
    cache = f["cache"].reshape(-1, 14)
    V = camera.col_major(cam.view)
    VP = camera.col_major(cam.view_proj)
    fx, fy = float(u[48]), float(u[49])
    W, H = cam.width, cam.height
    checked = 0
    for r in range(0, s.count, 37):
        w = cache[r]
        if w[7] == 0xFFFFFFFF:
            continue
        i = srt[r]
        mu = s.positions[i].astype(np.float64)
        clip = VP @ np.append(mu, 1.0)
        px = ((clip[0] / clip[3]) * 0.5 + 0.5) * W
        py = (1 - ((clip[1] / clip[3]) * 0.5 + 0.5)) * H
        assert abs(w[0:2].view(np.float32)[0] - px) < 1e-3 and abs(w[0:2].view(np.float32)[1] - py) < 1e-3
        vc = (V @ np.append(mu, 1.0))[:3]
        vz = min(vc[2], -0.001)
        J = np.array([[fx / vz, 0, -fx * vc[0] / vz ** 2], [0, fy / vz, -fy * vc[1] / vz ** 2]])
        cov2 = J @ V[:3, :3] @ _cov3(s.scales[i], s.rotations[i]) @ V[:3, :3].T @ J.T
        blur = cov2 + 0.3 * np.eye(2)
        conic = np.linalg.inv(blur)
        k = w[2:5].view(np.float32).astype(np.float64)
        # cache words 2-4 hold (inv[0][0], inv[0][1], inv[1][1]) of the blurred covariance
        np.testing.assert_allclose([k[0], k[1], k[2]], [conic[0, 0], conic[0, 1], conic[1, 1]], rtol=2e-3,
                                   atol=1e-7)
        # opacity: aa = sqrt(det(cov2)/det(blur)), clamp to 0.99, f16
        aa = math.sqrt(max(np.linalg.det(cov2) / np.linalg.det(blur), 0.0))
        op = np.float16(min(max(float(s.opacities[i]) * aa, 0.0), 0.99))
        got = np.uint16(w[6] >> 16).view(np.float16)
        assert abs(float(got) - float(op)) <= 2 * float(np.spacing(op)) + 1e-6
        # radius = ceil(3 sqrt(lambda_max)) of the blurred covariance
        lam = np.linalg.eigvalsh(blur).max()
        rad = float(np.uint16(w[6] & 0xFFFF).view(np.float16))
        assert abs(rad - math.ceil(3 * math.sqrt(lam))) <= 1.0
        checked += 1
    assert checked > 50


def test_sh_degree0_is_dc(small_frame):
    s, cam, plan, srt, u, f = small_frame
    cache = f["cache"].reshape(-1, 14)
    vis = cache[:, 7] != 0xFFFFFFFF
    rg = cache[vis, 9]
    r = (rg & 0xFFFF).astype(np.uint16).view(np.float16)
    g = (rg >> 16).astype(np.uint16).view(np.float16)
    dc = s.colors[srt[vis]]
    np.testing.assert_array_equal(r, dc[:, 0].astype(np.float16))
    np.testing.assert_array_equal(g, dc[:, 1].astype(np.float16))


def single_splat_scene(pos, logscale, opacity, color):
    n = 1
    s = scenes.Scene(np.asarray([pos], np.float32), np.asarray([logscale], np.float32),
                     np.asarray([[1, 0, 0, 0]], np.float32), np.asarray([opacity], np.float32),
                     np.asarray([color], np.float32), None, 0, scenes.pack_material(n))
    return s


def test_single_isotropic_splat_closed_form():
    s = single_splat_scene([0.0, 0.0, 0.0], [math.log(0.02)] * 3, 0.8, [0.9, 0.5, 0.1])
    cam = camera.Camera(camera.orbit_eye(3.0, 0.0, 0.0), (0, 0, 0), 525.0, 525.0, 64, 48)
    plan, srt, u = frame_inputs(s, cam)
    f = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_FP64)
    rgba = f["rgba"].reshape(48, 64, 4)
    w = f["cache"]
    cx, cy = w[0:2].view(np.float32)
    assert abs(cx - 32.0) < 1e-4 and abs(cy - 24.0) < 1e-4
    k0, k1, k2 = w[2:5].view(np.float32).astype(np.float64)
    assert abs(k1) < 1e-6 and abs(k0 - k2) < 1e-6 * k0  # radially symmetric footprint
    o = float(np.uint16(w[6] >> 16).view(np.float16))
    col = [float(np.float16(c)) for c in (0.9, 0.5, 0.1)]
    for (X, Y) in [(32, 24), (31, 24), (35, 20), (28, 27), (40, 24)]:
        dx, dy = X + 0.5 - cx, Y + 0.5 - cy
        p = -0.5 * k0 * dx * dx - k1 * dx * dy - 0.5 * k2 * dy * dy
        a = min(0.99, o * math.exp(p))
        valid = p > -4 and a > 1 / 255
        T = 1 - a if valid else 1.0
        expect = [(c * a if valid else 0.0) + T * bg for c, bg in zip(col, (0.02, 0.02, 0.04))] + [1 - T]
        np.testing.assert_allclose(rgba[Y, X], expect, atol=2e-6)
    # symmetry of the image about the splat centre (pixel centres at +-0.5)
    np.testing.assert_allclose(rgba[24, 32 - 3], rgba[24, 32 + 2], atol=1e-6)


def test_tile_loaded_ignores_pixels_outside_the_viewport():
    """1000 copies of a splat wide in x and about 7 px tall, centred on row 20 of a 64x40 image:
    rows 40-47 of the last tile row lie outside the viewport, where the power is below -4 and a
    pixel never saturates.  Only the viewport's pixels decide how many 256-entry batches a tile
    loads: float64 closed form of each pixel's saturation count, n = ceil(ln 0.001 / ln(1 - a))."""
    n = 1000
    sc = np.asarray([[math.log(2.0), math.log(0.04), math.log(2.0)]], np.float32)
    s = scenes.Scene(np.zeros((n, 3), np.float32), np.tile(sc, (n, 1)),
                     np.tile(np.asarray([[1, 0, 0, 0]], np.float32), (n, 1)), np.full(n, 0.99, np.float32),
                     np.full((n, 3), 0.5, np.float32), None, 0, scenes.pack_material(n))
    cam = camera.Camera(camera.orbit_eye(3.0, 0.0, 0.0), (0, 0, 0), 525.0, 525.0, 64, 40)
    plan, srt, u = frame_inputs(s, cam)
    f = oracle.run(s, srt, u, cam.width, cam.height)
    w = f["cache"][:14]
    cx, cy = w[0:2].view(np.float32).astype(np.float64)
    k0, k1, k2 = w[2:5].view(np.float32).astype(np.float64)
    o = float(np.uint16(w[6] >> 16).view(np.float16))
    assert abs(cy - 20.0) < 1e-4 and 6.5 < 1 / math.sqrt(k2) < 7.5 and k0 < 1e-4  # wide in x, ~7 px tall

    def sat(X, Y):
        dx, dy = X + 0.5 - cx, Y + 0.5 - cy
        p = -0.5 * k0 * dx * dx - k1 * dx * dy - 0.5 * k2 * dy * dy
        a = min(0.99, o * math.exp(p))
        if not (p > -4 and a > 1 / 255):
            return n + 1  # never
        return math.ceil(math.log(0.001) / math.log(1 - a))

    assert sat(0, 40) == n + 1 and sat(0, 39) <= n  # below the viewport: never; its last row: saturates
    for ty in range(plan["tileRows"]):
        for tx in range(plan["tileColumns"]):
            m = oracle.morton2(tx, ty)
            assert f["tile_counts"][m] == n
            last = max(sat(X, Y) for X in range(16 * tx, 16 * tx + 16) for Y in range(16 * ty, min(16 * ty + 16, 40)))
            assert f["tile_loaded"][m] == min(n, 256 * ((last - 1) // 256 + 1)) < n


@pytest.mark.parametrize("opacity,loaded", [(0.9, 256), (0.034, 256), (0.02, 512), (0.01, 768), (0.003, 1000)])
def test_tile_loaded_closed_form(opacity, loaded):
    """1000 copies of one wide splat over a 64x48 image: every pixel sees alpha = o * exp(p) with
    |p| < 0.01, so its transmittance (1 - alpha)^n drops below 0.001 at n = ln(0.001) / ln(1 - alpha)
    (3, 200, 344, 688 entries; never below 1/255 opacity).  The rasterizer loads whole 256-entry
    batches up to the one holding the last pixel's saturating entry, or the whole list."""
    n = 1000
    s = scenes.Scene(np.zeros((n, 3), np.float32), np.full((n, 3), math.log(2.0), np.float32),
                     np.tile(np.asarray([[1, 0, 0, 0]], np.float32), (n, 1)), np.full(n, opacity, np.float32),
                     np.full((n, 3), 0.5, np.float32), None, 0, scenes.pack_material(n))
    cam = camera.Camera(camera.orbit_eye(3.0, 0.0, 0.0), (0, 0, 0), 525.0, 525.0, 64, 48)
    plan, srt, u = frame_inputs(s, cam)
    f = oracle.run(s, srt, u, cam.width, cam.height)
    counts = f["tile_counts"]
    used = counts > 0
    assert used.sum() == 12 and np.all(counts[used] == n)
    np.testing.assert_array_equal(f["tile_loaded"][used], loaded)
    assert np.all(f["tile_loaded"][~used] == 0)
    a = float(np.float16(opacity))
    if a > 1 / 255:
        sat = math.ceil(math.log(0.001) / math.log(1 - a))
        assert loaded == min(n, 256 * ((sat - 1) // 256 + 1))


def test_bruteforce_blend_matches_oracle(small_frame):
    """Per pixel: walk the oracle's per-tile order front to back in float64 with the fp32 cache
    values and compare with the oracle's fp32 image."""
    s, cam, plan, srt, u, f = small_frame
    cache = f["cache"].reshape(-1, 14)
    rgba = f["rgba"].reshape(cam.height, cam.width, 4)
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(300):
        X, Y = int(rng.integers(0, cam.width)), int(rng.integers(0, cam.height))
        m = int(morton(X // 16, Y // 16))
        lo, c = f["tile_offsets"][m], f["tile_counts"][m]
        T, C = 1.0, np.zeros(3)
        for r in f["entries"][lo:lo + c]:
            w = cache[r]
            cx, cy = w[0:2].view(np.float32).astype(np.float64)
            k0, k1, k2 = w[2:5].view(np.float32).astype(np.float64)
            dx, dy = X + 0.5 - cx, Y + 0.5 - cy
            p = -0.5 * k0 * dx * dx - k1 * dx * dy - 0.5 * k2 * dy * dy
            o = float(np.uint16(w[6] >> 16).view(np.float16))
            a = min(0.99, o * math.exp(p))
            if p > -4 and a > 1 / 255 and T >= 0.001:
                col = np.array([np.uint16(w[9] & 0xFFFF).view(np.float16), np.uint16(w[9] >> 16).view(np.float16),
                                np.uint16(w[10] & 0xFFFF).view(np.float16)], np.float64)
                C += col * a * T
                T *= 1 - a
        exp = np.append(C + T * np.array([0.02, 0.02, 0.04]), 1 - T)
        worst = max(worst, float(np.abs(rgba[Y, X] - exp).max()))
    assert worst < 2e-3  # float64 vs fp32 rounding; threshold flips are rare at 300 pixels
    # the FP64 twin agrees with the fp32 oracle on nearly every pixel
    d = np.abs(f["rgba"] - f["rgba64"])
    assert np.mean(d) < 1e-5 and np.quantile(d, 0.999) < 1e-3


def test_raw_depth_order_diagnostic(dense_frame):
    # the 20-bit key order inverts the live reference's raw-bit order only inside q20 ties
    f = dense_frame[-1]
    assert f["raw_inversions"].sum() <= 0.01 * f.stats["refs"]


def test_f16_rounding_matches_numpy():
    """The oracle's binary16 rounding (every f16 operation of the f16 composite goes through it)
    against numpy's double -> half conversion: random values, exact ties, subnormals, overflow."""
    rng = np.random.default_rng(0)
    e = rng.integers(-14, 16, 20000)
    m = rng.integers(1024, 2048, 20000)
    x = np.concatenate([
        rng.uniform(-70000, 70000, 20000), rng.standard_normal(20000) * 1e-3, rng.standard_normal(20000) * 1e-6,
        np.ldexp(rng.integers(0, 2 ** 12, 20000).astype(np.float64) * 2 + 1, -25),  # subnormal midpoints
        np.ldexp(m + 0.5, e - 10),                                                  # normal midpoints
        [65504.0, 65519.99, 65520.0, -65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0, -0.0]])
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float64)
    got = oracle.f16_round(x)
    np.testing.assert_array_equal(got, want)
    # the f16 shader's constants 1h/255h, 0.001h, 0.02h, 0.04h
    np.testing.assert_array_equal(oracle.f16_round([1 / 255, 0.001, 0.02, 0.04]),
                                  [257 / 65536, 1049 / 2 ** 20, 1311 / 65536, 1311 / 32768])


@pytest.fixture(scope="module")
def small_frame_f16(small_frame):
    s, cam, plan, srt, u, _ = small_frame
    return oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_F16)


def test_f16_composite_half_walk(small_frame, small_frame_f16):
    """Per pixel: walk the tile's order with numpy float16 scalars for T, colour and weight
    (each numpy half operation rounds once) and float32 for the power / alpha, following
    gpu_tile_splat_composite_f16.wgsl:180-231 and :253-259; the oracle's f16 image must match
    bit for bit."""
    s, cam, plan, srt, u, f = small_frame
    g = small_frame_f16
    lib = oracle.load()
    cache = f["cache"].reshape(-1, 14)
    rgba = g["rgba"].reshape(cam.height, cam.width, 4)
    h, f32 = np.float16, np.float32
    a_th, t_cut = f32(h(1.0) / h(255.0)), h(0.001)
    bg = (h(0.02), h(0.02), h(0.04))
    rng = np.random.default_rng(1)
    blended = 0
    for _ in range(300):
        X, Y = int(rng.integers(0, cam.width)), int(rng.integers(0, cam.height))
        m = int(morton(X // 16, Y // 16))
        lo, c = f["tile_offsets"][m], f["tile_counts"][m]
        T, C = h(1.0), [h(0.0)] * 3
        p00x, p00y = f32((X & ~1) + 0.5), f32((Y & ~1) + 0.5)
        for r in f["entries"][lo:lo + c]:
            if T < t_cut:
                break
            w = cache[r]
            cx, cy = w[0:2].view(np.float32)
            k0, k1, k2 = w[2:5].view(np.float32)
            dx, dy = p00x - cx, p00y - cy
            if X & 1:
                dx = dx + f32(1.0)
            if Y & 1:
                dy = dy + f32(1.0)
            ca, cbx, cc = k0 * f32(-0.5), -k1, k2 * f32(-0.5)
            power = (ca * dx * dx + cbx * dx * dy) + cc * dy * dy
            o = f32(np.uint16(w[6] >> 16).view(np.float16))
            alpha = min(f32(0.99), o * f32(lib.orc_expf(float(power))))
            a16 = h(alpha)
            new_t = T * (h(1.0) - a16)
            valid = power > f32(-4.0) and alpha > a_th and f32(T) >= f32(t_cut)
            wgt = a16 * T * (h(1.0) if valid else h(0.0))
            col = (np.uint16(w[9] & 0xFFFF).view(np.float16), np.uint16(w[9] >> 16).view(np.float16),
                   np.uint16(w[10] & 0xFFFF).view(np.float16))
            C = [ci + cv * wgt for ci, cv in zip(C, col)]
            if valid:
                T = new_t
                blended += 1
        exp = [f32(ci + T * b) for ci, b in zip(C, bg)] + [f32(1.0) - f32(T)]
        np.testing.assert_array_equal(rgba[Y, X], np.array(exp, np.float32), err_msg=f"pixel {X},{Y}")
    assert blended > 300


def test_f16_composite_vs_fp32_and_fp64(small_frame, small_frame_f16):
    """The f16 image: same lists as fp32, rgb values binary16-representable, and within the
    error that binary16 accumulation allows of the fp32 image and the FP64 twin."""
    f, g = small_frame[-1], small_frame_f16
    np.testing.assert_array_equal(f["entries"], g["entries"])
    rgb = g["rgba"].reshape(-1, 4)[:, :3]
    np.testing.assert_array_equal(rgb, rgb.astype(np.float16).astype(np.float32))
    d32 = np.abs(g["rgba"] - f["rgba"])
    d64 = np.abs(g["rgba"].astype(np.float64) - f["rgba64"])
    # measured: mean 9.8e-5, max 3.3e-3 (fp32); max 5.1e-3 (FP64 twin)
    assert d32.mean() < 2e-4 and d32.max() < 0.01, (d32.mean(), d32.max())
    assert d64.mean() < 2e-4 and d64.max() < 0.01, (d64.mean(), d64.max())
    # depth votes gate on f32(w) > 0.01, so a pixel's voter set can change with the f16 weight
    dd = np.abs(g["out_depth"] - f["out_depth"])
    assert np.mean(dd > 1e-4) < 0.01


def test_quaternion_sign_flip_changes_nothing():
    # q and -q are one rotation: the rotation matrix is quadratic in q, so every product, and with
    # it every projection record and list, is the same bit for bit (the counterpart, for this
    # path's projection, of the covariance sign-flip invariance that
    # tests/renderer/conicProjection.test.mjs:17-33 checks on the reference's JS conic module)
    s = scenes.scene_config0(n=3000, seed=17)
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    a = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)
    s.rotations = -s.rotations
    b = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)
    for k in ("cache", "depth", "entries", "keys"):
        np.testing.assert_array_equal(a[k], b[k])


def test_footprint_radius_clamp():
    # the tile box uses min(radius, 8 tiles) (gpu_project_splats.wgsl:210-218; the clamp the
    # reference's lifecycle test requires, tests/renderer/computeCompositorLifecycle.test.mjs:48-64):
    # no box spans more than 17 tiles per axis, and interior splats with a larger radius span 16 or 17
    s = scenes.scene_config0(n=400, seed=23)
    s.scales[:] = np.float32(-0.5)  # exp(-0.5) * 3 sigma: radii of hundreds of pixels
    cam = camera.Camera((0.0, 0.0, 3.0), (0, 0, 0), 525.0, 525.0, 640, 480)
    plan, srt, u = frame_inputs(s, cam)
    f = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)
    c = f["cache"].reshape(-1, 14)
    vis = c[:, 7] != 0xFFFFFFFF
    x0, y0, x1, y1 = (c[vis, 7] & 0xFF), (c[vis, 7] >> 8) & 0xFF, (c[vis, 7] >> 16) & 0xFF, c[vis, 7] >> 24
    assert np.all(x1 - x0 + 1 <= 17) and np.all(y1 - y0 + 1 <= 17)
    px, py = c[vis, 0].view(np.float32), c[vis, 1].view(np.float32)
    rad = np.array([float(np.uint16(w & 0xFFFF).view(np.float16)) for w in c[vis, 6]])
    inner = (rad > 128) & (px > 200) & (px < 440) & (py > 200) & (py < 280)
    assert inner.sum() > 20
    assert np.all(np.isin(x1[inner] - x0[inner] + 1, (16, 17))) and np.all(np.isin(y1[inner] - y0[inner] + 1, (16, 17)))


def test_splat_scale_scales_the_covariance():
    # splatScale multiplies the three scales (uniform word 24): the projected covariance, recovered
    # from the conic as inv(conic) - 0.3 I, grows by splatScale^2
    s = scenes.scene_config0(n=2000, seed=29)
    cam = scenes.cameras_config0()[0]
    plan, srt, u = frame_inputs(s, cam)
    u2 = u.copy()
    u2[24] = np.float32(2.0) * u[24]
    a = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)["cache"].reshape(-1, 14)
    b = oracle.run(s, srt, u2, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)["cache"].reshape(-1, 14)
    both = (a[:, 7] != 0xFFFFFFFF) & (b[:, 7] != 0xFFFFFFFF)

    def cov(c):
        k = c[:, 2:5].view(np.float32).astype(np.float64)
        m = np.stack([np.stack([k[:, 0], k[:, 1]], -1), np.stack([k[:, 1], k[:, 2]], -1)], -2)
        return np.linalg.inv(m) - 0.3 * np.eye(2)

    ca, cb = cov(a[both]), cov(b[both])
    big = np.abs(ca[:, 0, 0]) > 1.0  # away from the 0.3 low-pass, where fp32 keeps the ratio
    assert big.sum() > 100
    np.testing.assert_allclose(cb[big, 0, 0] / ca[big, 0, 0], 4.0, rtol=2e-3)
    np.testing.assert_allclose(cb[big, 1, 1] / ca[big, 1, 1], 4.0, rtol=2e-3)
