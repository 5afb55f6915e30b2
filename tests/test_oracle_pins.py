"""Pins of the oracle functions that the restatement alone would leave open: the SH evaluation of
degrees 1-3 (src/shaders/gpu_project_splats.wgsl:293-351) and the alpha-weighted G-buffer votes of
the compositor (src/shaders/gpu_tile_splat_composite.wgsl:283-299,321-369).

SH: the oracle's projected colours (fp32, differentiable mode, so no f16 rounding) are read for
one coefficient slot at a time over a spherical product quadrature, with a DC of 0.5 that keeps
the clamp inactive.  The responses must be
  * orthonormal over the sphere (Gram matrix = identity to 1e-5): a wrong normalising constant,
    a wrong polynomial, a dropped or doubled term, or two slots reading one coefficient fails;
  * of their band's parity, Y(-d) = (-1)^l Y(d), which ties every slot to its band;
  * zero on their nodal sets (the coordinate planes and cones of the real SH basis in the
    reference's order) and of the sign the WGSL's positive constants give at a direction in
    general position, which fixes which basis function each slot holds;
  * equal to the textbook values sqrt(3 / 4pi) and sqrt(15 / 4pi) / 2 on the axes;
  * per channel: a coefficient in one colour channel moves that channel only.

Votes: closed forms for one and two splats over one pixel (the w > 0.01 gate, the weighted mean
depth, the renormalised oct normal, the emissive sum weighted by w rather than the gated weight),
and a float64 brute-force walk of the votes over random pixels of a scene with materials,
normals and emissive, against the oracle's G-buffer words.
"""
import math

import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import frame_inputs

SQRT = math.sqrt
PI4 = 4.0 * math.pi


def _quadrature(nz=10, nphi=20):
    """Gauss-Legendre in z times a uniform grid in phi: exact for polynomials of degree <= 19 in
    (x, y, z) on the sphere; weights sum to 4 pi.  nphi even, so the set is closed under d -> -d."""
    z, wz = np.polynomial.legendre.leggauss(nz)
    phi = (np.arange(nphi) + 0.5) * (2 * math.pi / nphi)
    Z, P = np.meshgrid(z, phi, indexing="ij")
    W = np.repeat(wz, nphi) * (2 * math.pi / nphi)
    r = np.sqrt(1 - Z.ravel() ** 2)
    d = np.stack([r * np.cos(P.ravel()), r * np.sin(P.ravel()), Z.ravel()], 1)
    return d, W


def _sh_responses(dirs, amp=0.25, slots=range(1, 16)):
    """Oracle colour - DC per (slot, direction) for channels (r, g, b) with +amp in r, 0 in g and
    -amp in b on that slot: shape (len(slots), len(dirs), 3), in units of amp."""
    dirs = np.asarray(dirs, np.float32)
    slots = list(slots)
    m, k = len(dirs), 15
    n = m * len(slots)
    pos = np.tile(-dirs, (len(slots), 1))  # camera at the origin: view dir = cam - centre = d
    sh = np.zeros((n, k, 3), np.float32)
    for i, j in enumerate(slots):
        sh[i * m:(i + 1) * m, j - 1, 0] = amp
        sh[i * m:(i + 1) * m, j - 1, 2] = -amp
    s = scenes.Scene(pos, np.full((n, 3), -3.0, np.float32), np.tile(np.float32([1, 0, 0, 0]), (n, 1)),
                     np.zeros(n, np.float32), np.full((n, 3), 0.5, np.float32), sh, 3, scenes.pack_material(n))
    plan = compositor.planTileSplatCompositor({"viewportWidth": 64, "viewportHeight": 64, "tileSizePx": 16,
                                               "splatCount": n})
    vp = np.zeros(16, np.float32)
    vp[14], vp[15] = 0.5, 1.0  # clip = (0, 0, 0.5, 1) for every centre: all visible
    proj = np.zeros(16, np.float32)
    proj[0] = proj[5] = 1.0
    u = camera.write_frame_uniforms(vp, plan, 1.0, 3, [0.0, 0.0, 0.0], np.eye(4, dtype=np.float32).ravel(), proj)
    f = oracle.run(s, np.arange(n, dtype=np.uint32), u, 64, 64, oracle.FLAG_DIFF | oracle.FLAG_NO_COMPOSITE)
    rec = f["diffrec"].reshape(n, 4).astype(np.float64)
    assert np.all(rec[:, 1:] > 0.05) and np.all(rec[:, 1:] < 0.95)  # the [0, 1] clamp never acts
    rgb = (rec[:, 1:] - 0.5) / amp
    return rgb.reshape(len(slots), m, 3)


BAND = [1] * 3 + [2] * 5 + [3] * 7  # slot 1..15 -> band


@pytest.fixture(scope="module")
def sh_quadrature():
    d, w = _quadrature()
    return d, w, _sh_responses(d)


def test_sh_basis_orthonormal(sh_quadrature):
    d, w, Y = sh_quadrature
    r = Y[:, :, 0]
    G = (r * w) @ r.T
    np.testing.assert_allclose(G, np.eye(15), atol=1e-5)
    # g untouched, b the negated response: one channel per coefficient
    np.testing.assert_array_equal(Y[:, :, 1], 0.0)
    np.testing.assert_allclose(Y[:, :, 2], -Y[:, :, 0], atol=1e-6)


def test_sh_band_parity(sh_quadrature):
    d, w, Y = sh_quadrature
    # the antipode of every quadrature direction is in the set
    idx = np.array([int(np.argmin(np.abs(d + x).sum(1))) for x in d])
    assert np.abs(d[idx] + d).max() < 1e-12
    for j in range(15):
        sign = (-1) ** BAND[j]
        np.testing.assert_allclose(Y[j, idx, 0], sign * Y[j, :, 0], atol=2e-6, err_msg=f"slot {j + 1}")


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# Nodal sets of the real SH basis in the reference's slot order (gpu_project_splats.wgsl:300-347):
# each slot lists directions where its function vanishes.  A coordinate that is exactly zero
# stays exactly zero through the oracle's normalisation, so those are required to be 0 exactly;
# the cones (z^2 = 1/3, 1/5, 3/5, |x| = |y|, y = sqrt3 x, x = sqrt3 y) only to rounding.
_A, _B = 0.48, 0.6
NODAL = {
    1: [(1, 0, 0), (0, 0, 1), (_A, 0, _B)],                      # y
    2: [(1, 0, 0), (0, 1, 0), (_A, _B, 0)],                      # z
    3: [(0, 1, 0), (0, 0, 1), (0, _A, _B)],                      # x
    4: [(0, 0, 1), (0, _A, _B), (_A, 0, _B)],                    # xy
    5: [(1, 0, 0), (_A, 0, _B), (_A, _B, 0)],                    # yz
    6: [(SQRT(2), 0, 1), (1, 1, 1), (0, SQRT(2), 1)],            # 3z^2 - 1
    7: [(0, 1, 0), (_A, _B, 0), (0, _A, _B)],                    # xz
    8: [(1, 1, 0), (1, -1, 0.7), (-1, 1, 0.3)],                  # x^2 - y^2
    9: [(1, 0, 0), (1, SQRT(3), 0.2), (-1, SQRT(3), 0.5)],       # y (3x^2 - y^2)
    10: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],                       # xyz
    11: [(1, 0, 0), (0.3, 0.4, SQRT(0.25 / 4)), (_A, 0, _B)],    # y (5z^2 - 1)
    12: [(1, 0, 0), (0, 1, 0), (SQRT(0.4), 0, SQRT(0.6))],       # z (5z^2 - 3)
    13: [(0, 1, 0), (0.3, 0.4, SQRT(0.25 / 4)), (0, _A, _B)],    # x (5z^2 - 1)
    14: [(1, 0, 0), (1, 1, 0.5), (1, -1, -0.4)],                 # z (x^2 - y^2)
    15: [(0, 1, 0), (SQRT(3), 1, 0.2), (SQRT(3), -1, 0.4)],      # x (x^2 - 3y^2)
}


def test_sh_nodal_sets_signs_and_axis_values():
    dirs = sorted({_unit(v).tobytes(): _unit(v) for vs in NODAL.values() for v in vs}.items())
    probe = _unit((0.48, 0.6, 0.64))                       # general position
    axes = [(0, 1, 0), (0, 0, 1), (1, 0, 0), _unit((1, 1, 0))]
    all_dirs = [v for _, v in dirs] + [probe] + [_unit(a) for a in axes]
    Y = _sh_responses(all_dirs)[:, :, 0]
    key = {k: i for i, (k, _) in enumerate(dirs)}
    for slot, vs in NODAL.items():
        for v in vs:
            u = _unit(v)
            y = Y[slot - 1, key[u.tobytes()]]
            assert (y == 0.0) if _zero_by_coordinate(slot, u) else abs(y) < 3e-6, (slot, v, y)
    # away from the nodal sets every slot is nonzero, with the WGSL's positive-constant signs
    x, y_, z = probe
    sign = np.sign([y_, z, x, x * y_, y_ * z, 3 * z * z - 1, x * z, x * x - y_ * y_, y_ * (3 * x * x - y_ * y_),
                    x * y_ * z, y_ * (5 * z * z - 1), z * (5 * z * z - 3), x * (5 * z * z - 1), z * (x * x - y_ * y_),
                    x * (x * x - 3 * y_ * y_)])
    got = Y[:, len(dirs)]
    assert np.all(np.abs(got) > 1e-3)
    np.testing.assert_array_equal(np.sign(got), sign)
    # textbook normalisation on the axes: Y_1 = sqrt(3/4pi) at its axis, Y_2 (xy) = sqrt(15/4pi)/2
    # at (1, 1, 0)/sqrt2, Y_2 (3z^2-1) = sqrt(5/16pi) * 2 at the pole
    i0 = len(dirs) + 1
    c1 = SQRT(3 / PI4)
    np.testing.assert_allclose([Y[0, i0], Y[1, i0 + 1], Y[2, i0 + 2]], [c1, c1, c1], rtol=2e-6)
    np.testing.assert_allclose(Y[3, i0 + 3], SQRT(15 / PI4) / 2, rtol=2e-6)
    np.testing.assert_allclose(Y[5, i0 + 1], SQRT(5 / (16 * math.pi)) * 2, rtol=2e-6)


def _zero_by_coordinate(slot, u):
    """True when the slot's function vanishes at u because a coordinate is exactly zero."""
    x, y, z = u
    return {1: y == 0, 2: z == 0, 3: x == 0, 4: x == 0 or y == 0, 5: y == 0 or z == 0, 7: x == 0 or z == 0,
            9: y == 0, 10: x == 0 or y == 0 or z == 0, 11: y == 0, 12: z == 0, 13: x == 0, 14: z == 0,
            15: x == 0}.get(slot, False)


# ---- G-buffer votes ----------------------------------------------------------------------


def _vote_scene(n, pos, logscale, opacity, rough, metal, normals, emissive):
    s = scenes.Scene(np.asarray(pos, np.float32), np.asarray(logscale, np.float32),
                     np.tile(np.float32([1, 0, 0, 0]), (n, 1)), np.asarray(opacity, np.float32),
                     np.full((n, 3), 0.5, np.float32), None, 0,
                     scenes.pack_material(n, np.asarray(rough, np.float32), np.asarray(metal, np.float32)),
                     np.asarray(normals, np.float32).ravel(), np.asarray(emissive, np.float32))
    return s


def _h(word, hi=False):
    return float(np.uint16((int(word) >> 16) if hi else (int(word) & 0xFFFF)).view(np.float16))


def _oct_decode(word):
    ex, ey = _h(word), _h(word, True)
    n = np.array([ex, ey, 1 - abs(ex) - abs(ey)])
    if n[2] < 0:
        n[0], n[1] = (1 - abs(ey)) * math.copysign(1, ex), (1 - abs(ex)) * math.copysign(1, ey)
    return n / np.linalg.norm(n)


def _votes64(cache, depth_bits, entries, X, Y, gbuffer=True):
    """Float64 walk of one pixel's list: (depth, normal, (rough, metal), emissive, confidence)."""
    T = 1.0
    gdw = gws = 0.0
    acc_n = np.zeros(3)
    mat = np.zeros(2)
    em = np.zeros(3)
    conf = 0.0
    for r in entries:
        w = cache[r]
        cx, cy = w[0:2].view(np.float32).astype(np.float64)
        k0, k1, k2 = w[2:5].view(np.float32).astype(np.float64)
        dx, dy = X + 0.5 - cx, Y + 0.5 - cy
        p = -0.5 * k0 * dx * dx - k1 * dx * dy - 0.5 * k2 * dy * dy
        a = min(0.99, _h(w[6], True) * math.exp(p))
        if not (p > -4 and a > 1 / 255 and T >= 0.001):
            continue
        wt = a * T
        T *= 1 - a
        gw = wt if wt > 0.01 else 0.0
        gdw += gw * float(np.uint32(depth_bits[r]).view(np.float32))
        gws += gw
        acc_n += gw * _oct_decode(w[8])
        mat += gw * np.array([_h(w[5]), _h(w[5], True)])
        em += wt * np.array([_h(w[11]), _h(w[11], True), _h(w[12])])
        conf += gw * _h(w[13])
    safe = max(gws, 1e-4)
    v = acc_n + np.array([0.0, 1e-4, 0.0])
    return gdw / safe, v / np.linalg.norm(v), mat / safe, em, conf / safe


def _read_votes(f, W, X, Y):
    p = Y * W + X
    gm = f["gb_material"].reshape(-1, 4)[p]
    return (float(f["out_depth"][p]), _oct_decode(f["gb_normal"][p]), np.array([_h(gm[0]), _h(gm[0], True)]),
            np.array([_h(gm[1]), _h(gm[1], True), _h(gm[2])]), _h(gm[3]))


def _f16_close(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= 2.0 ** -10 * np.maximum(np.abs(b), 2.0 ** -14) + 1e-7)


def test_votes_one_and_two_splats_closed_form():
    """Two splats on one line of sight, front splat first: at the centre pixel the front weight is
    alpha1, the back weight alpha2 (1 - alpha1).  With both above 0.01 the depth is their weighted
    mean; the normal the renormalised weighted sum of the two unit normals; roughness / metalness /
    confidence weighted means; emissive the w-weighted sum.  With the back weight below 0.01 the
    back splat drops out of the gated votes but still adds its emissive."""
    cam = camera.Camera((0.0, 0.0, 3.0), (0, 0, 0), 400.0, 400.0, 64, 64)
    for o2, gated in [(0.6, False), (0.004 / (1 - 0.7) + 0.004, True)]:
        n = 2
        pos = [[0, 0, 0.2], [0, 0, -0.3]]                         # first splat nearer the camera
        normals = np.concatenate([[_unit((0, 0.6, 0.8)), _unit((0.8, 0, 0.6))]])
        s = _vote_scene(n, pos, [[math.log(0.05)] * 3] * 2, [0.7, o2], [0.25, 0.75], [0.5, 0.125], normals,
                        [[0.5, 0.25, 0.125], [1.0, 2.0, 4.0]])
        plan, srt, u = frame_inputs(s, cam)
        assert list(srt) == [1, 0]                               # back to front: far splat first
        f = oracle.run(s, srt, u, 64, 64, oracle.FLAG_GBUFFER)
        cache = f["cache"].reshape(-1, 14)
        X = Y = 32
        m = 12  # morton(2, 2): the tile of pixel (32, 32)
        lo, c = int(f["tile_offsets"][m]), int(f["tile_counts"][m])
        assert c == 2
        # ranks: 0 = the far splat, 1 = the near one (back-to-front sortedIndices); the tile's
        # front-to-back order puts the near one (smaller depthNdc) first
        front, back = (int(r) for r in f["entries"][lo:lo + 2])
        assert (front, back) == (1, 0)
        d, nrm, mat, em, conf = _read_votes(f, 64, X, Y)
        wts = []
        T = 1.0
        for r in (front, back):
            w = cache[r]
            cx, cy = w[0:2].view(np.float32).astype(np.float64)
            k0, k1, k2 = w[2:5].view(np.float32).astype(np.float64)
            dx, dy = X + 0.5 - cx, Y + 0.5 - cy
            a = min(0.99, _h(w[6], True) * math.exp(-0.5 * k0 * dx * dx - k1 * dx * dy - 0.5 * k2 * dy * dy))
            wts.append(a * T)
            T *= 1 - a
        gw = [x if x > 0.01 else 0.0 for x in wts]
        assert (gw[1] == 0.0) == gated
        dep = [float(np.uint32(f["depth"][r]).view(np.float32)) for r in (front, back)]
        want_d = (gw[0] * dep[0] + gw[1] * dep[1]) / max(sum(gw), 1e-4)
        assert abs(d - want_d) <= 2e-7 * abs(want_d)
        if gated:
            assert abs(d - dep[0]) <= 2e-7 * abs(dep[0])
        nv = gw[0] * _oct_decode(cache[front, 8]) + gw[1] * _oct_decode(cache[back, 8]) + [0, 1e-4, 0]
        assert np.arccos(min(1.0, float(nrm @ (nv / np.linalg.norm(nv))))) < 2e-3
        want_mat = (gw[0] * np.array([_h(cache[front, 5]), _h(cache[front, 5], True)]) +
                    gw[1] * np.array([_h(cache[back, 5]), _h(cache[back, 5], True)])) / max(sum(gw), 1e-4)
        assert _f16_close(mat, want_mat)
        want_em = sum(wt * np.array([_h(cache[r, 11]), _h(cache[r, 11], True), _h(cache[r, 12])])
                      for wt, r in zip(wts, (front, back)))
        assert _f16_close(em, want_em)
        want_conf = (gw[0] * _h(cache[front, 13]) + gw[1] * _h(cache[back, 13])) / max(sum(gw), 1e-4)
        assert _f16_close(conf, want_conf)


def test_votes_bruteforce_walk():
    """Random pixels of a scene with per-splat roughness / metalness, baked normals and emissive:
    the float64 walk of the votes over the oracle's own order equals the oracle's G-buffer."""
    rng = np.random.default_rng(31)
    n = 3000
    base = scenes.scene_config0(n, seed=31)
    normals = rng.normal(0, 1, (n, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    s = _vote_scene(n, base.positions, base.scales, base.opacities, rng.uniform(0, 1, n), rng.uniform(0, 1, n),
                    normals, rng.uniform(0, 2, (n, 3)))
    cam = camera.Camera(camera.orbit_eye(2.2, 0.3, 0.2), (0, 0, 0), 300.0, 300.0, 160, 128)
    plan, srt, u = frame_inputs(s, cam)
    f = oracle.run(s, srt, u, 160, 128, oracle.FLAG_GBUFFER)
    cache = f["cache"].reshape(-1, 14)

    def morton(x, y):
        return sum((((x >> b) & 1) << (2 * b)) | (((y >> b) & 1) << (2 * b + 1)) for b in range(8))

    checked = 0
    for _ in range(250):
        X, Y = int(rng.integers(0, 160)), int(rng.integers(0, 128))
        m = morton(X // 16, Y // 16)
        lo, c = int(f["tile_offsets"][m]), int(f["tile_counts"][m])
        want = _votes64(cache, f["depth"], f["entries"][lo:lo + c], X, Y)
        got = _read_votes(f, 160, X, Y)
        if want[0] == 0.0:
            assert got[0] == 0.0
            continue
        checked += 1
        assert abs(got[0] - want[0]) <= 1e-5 * abs(want[0]) + 1e-7, (X, Y)
        assert np.arccos(min(1.0, float(got[1] @ want[1]))) < 3e-3, (X, Y)
        assert _f16_close(got[2], want[2]) or np.abs(got[2] - want[2]).max() < 2e-3, (X, Y)
        assert np.abs(got[3] - want[3]).max() <= 2e-3 * max(1.0, np.abs(want[3]).max()), (X, Y)
        assert abs(got[4] - want[4]) < 2e-3, (X, Y)
    assert checked > 100
