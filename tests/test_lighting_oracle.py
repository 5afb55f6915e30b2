"""The deferred-lighting pass (src/shaders/gpu_deferred_lighting.wgsl, path without an environment
map) in the oracle, and its host-side inputs in the library, pinned by what the mathematics and the
reference's own code fix:

  * the 176-byte Params block, word by word from createDeferredLightingPass().encode
    (src/splatRenderer.ts:505-523);
  * the material-curve tables: generateLUT (src/materialCurves.ts:31-42) through uploadLUT's
    truncating binary16 conversion (:45-57) - never rounded up, exact where the value is a
    binary16 number, the identity / flat / saturated curves;
  * GGX: the distribution integrates to 1 over the hemisphere (int D(h) (n.h) dw = 1) for every
    roughness - a dropped pi or a roughness used where its square belongs fails;
  * Smith-Schlick with the direct-light k = (r + 1)^2 / 8 (Karis 2013), not the image-based
    k = r^2 / 2: G(1, 1) = 1, G(0, x) = 0, symmetric, the textbook value at r = 0.6;
  * Schlick's Fresnel: F(1) = F0, F(0) = 1, F(0.5) = F0 + (1 - F0) / 32;
  * closed forms of the whole pass: normal incidence (N = V = L) specular-only gives
    D(1) F0 / (4 + 1e-4) tone mapped; an ambient-only frame gives the ambient term alone;
  * invariances: a light below the horizon equals no light, a flipped stored normal changes
    nothing (faceForward), an emissive above the threshold adds exactly emissive * intensity
    after tone mapping, a background pixel passes its colour with alpha 0;
  * the bloom sample: a constant field acts as that AO lift; a field linear across columns is
    reproduced exactly by the bilinear sample inside the image, the left column wraps (repeat)
    and the top row clamps.
"""
import math

import numpy as np
import pytest

from hdsr import _lib
from oracle import oracle

def f16(x):
    return np.float32(np.float16(np.float32(x)))


def half_bits_to_f32(h):
    return np.asarray(h, np.uint16).view(np.float16).astype(np.float32)


def pack2(a, b):
    return int(np.float16(a).view(np.uint16)) | (int(np.float16(b).view(np.uint16)) << 16)


def identity_luts():
    return np.concatenate([lut_of(1.0, 0.0, 1.0)] * 3)


def lut_of(contrast, brightness, gamma):
    out = np.zeros(256, np.float32)
    lib = _lib.load()
    lib.hdsr_material_curve_lut(_lib.MaterialCurve(contrast, brightness, gamma), out.ctypes.data)
    return out


def block(vw, vh, inv16, cam3, light_dir, intensity, ambient, specular_only=0, em_int=3.0, em_thr=0.05,
          exposure=1.0):
    """The Params block laid out by hand from src/splatRenderer.ts:505-523."""
    p = np.zeros(44, np.float32)
    p[0], p[1], p[2], p[3] = vw, vh, 0.65, 0.0
    p[4:20] = inv16
    p[20:23] = cam3
    p[24:27] = light_dir
    p[28:31] = (1.0, 0.95, 0.9)
    p[31] = intensity
    p[32], p[33], p[34] = ambient, ambient, np.float32(ambient * 1.1)
    p[35] = 1.0 if specular_only else 0.0
    p[36], p[37], p[38], p[39], p[40] = em_int, em_thr, 1.0, 0.0, exposure
    return p


def test_params_block_layout_word_by_word():
    lib = _lib.load()
    p = _lib.LightingParams()
    lib.hdsr_lighting_params_default(C_byref(p))
    p.light_direction[:] = [0.1, -0.7, 0.3]
    p.light_intensity = 2.5
    p.ambient_intensity = 0.1
    p.specular_only = 1
    p.emissive_intensity = 4.0
    p.emissive_threshold = 0.2
    p.env_intensity = 0.7
    p.env_rotation = 1.25
    p.exposure = 1.5
    inv = np.arange(16, dtype=np.float32) * 0.5 - 3.0
    cam = np.array([1.0, 2.0, 3.0], np.float32)
    out = np.full(44, np.nan, np.float32)
    lib.hdsr_write_lighting_params(C_byref(p), inv.ctypes.data, cam.ctypes.data, 640, 360, out.ctypes.data)
    want = np.zeros(44, np.float32)
    want[0], want[1] = 640, 360          # viewport
    want[2], want[3] = 0.65, 0.0         # fallback roughness / metallic
    want[4:20] = inv                     # viewProjInverse
    want[20:23] = cam                    # cameraPos, [23] pad
    want[24:27] = (0.1, -0.7, 0.3)       # lightDirection, [27] pad
    want[28:31] = (1.0, 0.95, 0.9)       # lightColor (warm white)
    want[31] = 2.5                       # lightIntensity
    want[32], want[33] = 0.1, 0.1        # ambient
    want[34] = np.float32(0.1 * 1.1)     # ambient * 1.1, evaluated in double then stored
    want[35] = 1.0                       # specularOnly
    want[36], want[37], want[38], want[39], want[40] = 4.0, 0.2, 0.7, 1.25, 1.5
    np.testing.assert_array_equal(out.view(np.uint32), want.view(np.uint32))
    # the defaults: the reference viewer's fixed light, intensity 3, ambient 0.12 (src/main.ts:145-149)
    d = _lib.LightingParams()
    lib.hdsr_lighting_params_default(C_byref(d))
    ce = math.cos(0.6)
    assert list(d.light_direction) == [math.cos(0.8) * ce, -math.sin(0.6), math.sin(0.8) * ce]
    assert (d.light_intensity, d.ambient_intensity, d.emissive_intensity, d.emissive_threshold, d.exposure) == \
        (3.0, 0.12, 3.0, 0.05, 1.0)


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


def trunc_f16(v32):
    """uploadLUT's f32ToF16 (src/materialCurves.ts:45-57) written over the bit fields."""
    bits = np.float32(v32).view(np.uint32).item()
    sign = (bits >> 16) & 0x8000
    e = ((bits >> 23) & 0xFF) - 127 + 15
    frac = bits & 0x7FFFFF
    if e <= 0:
        h = sign
    elif e >= 31:
        h = sign | 0x7C00
    else:
        h = sign | (e << 10) | (frac >> 13)
    return half_bits_to_f32(np.uint16(h)).item()


def test_material_curve_tables():
    ident = lut_of(1.0, 0.0, 1.0)
    for i in range(256):
        v = np.float32(i / 255)
        assert ident[i] == trunc_f16(v)
        assert ident[i] <= v and v - ident[i] < 2.0 ** -10 * max(v, 2.0 ** -14)  # truncation: never up
    assert ident[0] == 0.0 and ident[255] == 1.0  # binary16 numbers stay exact; 0.2 and 0.4 truncate
    assert ident[51] < np.float32(0.2) and ident[102] < np.float32(0.4)
    assert np.all(lut_of(0.0, 0.0, 1.0) == 0.5)                   # flat curve
    assert np.all(lut_of(1.0, 1.0, 1.0)[128:] == 1.0)             # brightness saturates
    sq = lut_of(1.0, 0.0, 2.0)
    for i in (0, 64, 128, 200, 255):
        assert sq[i] == trunc_f16(np.float32((i / 255) ** 2))
    # a round-to-nearest conversion would differ from the truncation somewhere in the identity table
    assert any(float(np.float16(np.float32(i / 255))) != ident[i] for i in range(256))


def test_ggx_distribution_is_normalised():
    lib = oracle.load()
    for r in (0.1, 0.25, 0.5, 0.8, 1.0):
        # 2 pi int_0^1 D(mu) mu dmu, with t = 1 - mu log-spaced to resolve the peak at mu = 1
        t = np.concatenate([[0.0], np.logspace(-12, 0, 200001)])
        mu = 1.0 - t
        d = np.array([lib.orc_ggx_d(float(m), r) for m in mu[::50]], np.float64)
        mu_s = mu[::50]
        integral = 2 * math.pi * np.trapezoid(d * mu_s, -mu_s) if hasattr(np, "trapezoid") else \
            2 * math.pi * np.trapz(d * mu_s, -mu_s)
        assert abs(integral - 1.0) < 2e-3, (r, integral)
    # the peak value 1 / (pi a^2), a = r^2; fp32 evaluates d = (a^2 - 1) + 1, which cancels to
    # within one ulp of 1, a relative error of up to 2^-24 / a^2 in d (twice that in D)
    for r in (0.2, 0.7):
        a2 = r ** 4
        assert lib.orc_ggx_d(1.0, r) == pytest.approx(1.0 / (math.pi * a2), rel=2 * 2.0 ** -24 / a2 + 1e-6)


def test_smith_g_direct_light_k():
    lib = oracle.load()
    for r in (0.04, 0.3, 0.6, 1.0):
        assert lib.orc_smith_g(1.0, 1.0, r) == pytest.approx(1.0, abs=1e-7)
        assert lib.orc_smith_g(0.0, 0.7, r) == 0.0
        for a, b in ((0.2, 0.9), (0.5, 0.3)):
            assert lib.orc_smith_g(a, b, r) == lib.orc_smith_g(b, a, r)
    k = 1.6 ** 2 / 8  # Karis' k for direct light at r = 0.6 (0.32); the image-based k would be 0.18
    for x in (0.1, 0.4, 0.9):
        assert lib.orc_smith_g(x, 1.0, 0.6) == pytest.approx(x / (x * (1 - k) + k), rel=1e-6)


def test_schlick_fresnel():
    lib = oracle.load()
    for f0 in (0.04, 0.5, 0.95):
        assert lib.orc_fresnel(1.0, f0) == np.float32(f0)
        assert lib.orc_fresnel(0.0, f0) == pytest.approx(1.0, abs=1e-6)
        assert lib.orc_fresnel(0.5, f0) == pytest.approx(f0 + (1 - f0) / 32, rel=1e-6)


def gbuffer(W, H, color=(0.6, 0.5, 0.4, 0.9), depth=0.5, normal=(0.0, 0.0, 1.0), rough=0.5, metal=0.0,
            em=(0.0, 0.0, 0.0)):
    rgba = np.tile(np.asarray(color, np.float32), (H, W, 1))
    dep = np.full((H, W), depth, np.float32)
    nw = np.full((H, W), oracle.normals_to_oct_words(np.asarray(normal, np.float64)), np.uint32)
    mat = np.zeros((H, W, 4), np.uint32)
    mat[..., 0] = pack2(rough, metal)
    mat[..., 1] = pack2(em[0], em[1])
    mat[..., 2] = pack2(em[2], 0.0)
    return rgba, dep, nw, mat


EYE16 = np.eye(4, dtype=np.float32).reshape(16)  # inverse view-projection = identity: world = (ndc, depth)


def srgb_tonemap(c, exposure=1.0):
    e = np.float64(c) * exposure
    return (e / (e + 1.0)) ** (1.0 / 2.2)


def test_normal_incidence_specular_closed_form():
    # 2 x 2 image, pixel (0, 0): ndc (-0.5, 0.5), world (-0.5, 0.5, d); camera straight above it;
    # stored normal +z, light travelling along -z: N = V = L = H, every cosine 1
    W = H = 2
    d = 0.5
    rgba, dep, nw, mat = gbuffer(W, H, depth=d, rough=0.5, metal=0.0)
    cam = (-0.5, 0.5, d + 2.0)
    for r in (0.3, 0.5, 0.9):
        mat[..., 0] = pack2(r, 0.0)
        P = block(W, H, EYE16, cam, (0.0, 0.0, -1.0), 2.0, 0.1, specular_only=1)
        out = half_bits_to_f32(oracle.deferred_lighting(rgba, dep, nw, mat, P, identity_luts()))
        rl = lut_of(1.0, 0.0, 1.0)[int(np.float32(f16(r)) * 255)]
        spec = (1.0 / (math.pi * rl ** 4)) * 0.04 / (4.0 + 1e-4)  # D(1) G(1,1) F(1) / (4 N.V N.L + 1e-4)
        want = [srgb_tonemap(spec * lc * 2.0) for lc in (1.0, 0.95, 0.9)]
        np.testing.assert_allclose(out[0, 0, :3], want, rtol=2e-3)
        assert out[0, 0, 3] == f16(0.9)


def ambient_only(W, H, rgba, dep, nw, mat, ambient=0.2, exposure=1.0, ao=None, bloom=None, light=(0.0, 0.0, -1.0),
                 em_int=3.0):
    P = block(W, H, EYE16, (0.0, 0.0, 3.0), light, 0.0, ambient, em_int=em_int, exposure=exposure)
    return oracle.deferred_lighting(rgba, dep, nw, mat, P, identity_luts(), ao=ao, bloom=bloom)


def test_ambient_only_closed_form_and_invariances():
    W, H = 4, 3
    col = (0.6, 0.5, 0.4, 0.8)
    rgba, dep, nw, mat = gbuffer(W, H, color=col)
    out = half_bits_to_f32(ambient_only(W, H, rgba, dep, nw, mat, ambient=0.2, exposure=1.3))
    c = np.array([f16(v) for v in col[:3]], np.float64)
    lum = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    scale = lut_of(1.0, 0.0, 1.0)[int(np.float32(lum) * 255)] / lum
    alb = (c * scale) ** 2.2
    amb = np.array([0.2, 0.2, 0.22])
    want = srgb_tonemap(amb * alb, 1.3)
    for px in out.reshape(-1, 4):
        np.testing.assert_allclose(px[:3], want, rtol=2e-3)
        assert px[3] == f16(col[3])
    # no light: neither the normal nor the light direction matters
    base = ambient_only(W, H, rgba, dep, nw, mat)
    for n in ((0.3, 0.4, 0.86), (-1.0, 0.0, 0.0)):
        _, _, nw2, _ = gbuffer(W, H, normal=n)
        np.testing.assert_array_equal(ambient_only(W, H, rgba, dep, nw2, mat), base)
    np.testing.assert_array_equal(ambient_only(W, H, rgba, dep, nw, mat, light=(0.5, 0.2, 0.8)), base)


def lit(W, H, g, light, intensity=3.0, **kw):
    P = block(W, H, EYE16, (0.3, -0.2, 3.0), light, intensity, 0.15, **kw)
    return oracle.deferred_lighting(*g, P, identity_luts())


def test_light_below_horizon_face_forward_and_emissive():
    W, H = 5, 4
    g = gbuffer(W, H, normal=(0.0, 0.0, 1.0), rough=0.4, metal=0.3)
    # a light from below the surface (travelling +z, so L = -z) adds nothing
    np.testing.assert_array_equal(lit(W, H, g, (0.0, 0.0, 1.0)), lit(W, H, g, (0.0, 0.0, 1.0), intensity=0.0))
    # faceForward: the stored normal's sign does not matter
    g_flip = gbuffer(W, H, normal=(0.0, 0.0, -1.0), rough=0.4, metal=0.3)
    a = lit(W, H, g, (0.2, 0.3, -0.9))
    np.testing.assert_array_equal(lit(W, H, g_flip, (0.2, 0.3, -0.9)), a)
    assert np.any(a != lit(W, H, g, (0.2, 0.3, -0.9), intensity=0.0))  # the light does contribute
    # an emissive above the threshold adds emissive * intensity after tone mapping (AO is 1 here,
    # so its lift by the emissive changes nothing)
    em = (0.5, 0.25, 0.125)
    g_em = gbuffer(W, H, normal=(0.0, 0.0, 1.0), rough=0.4, metal=0.3, em=em)
    b = half_bits_to_f32(lit(W, H, g_em, (0.2, 0.3, -0.9), em_int=0.5))
    a32 = half_bits_to_f32(lit(W, H, g, (0.2, 0.3, -0.9), em_int=0.5))
    np.testing.assert_allclose(b[..., :3] - a32[..., :3], np.broadcast_to(np.array(em) * 0.5, b[..., :3].shape),
                               atol=2e-3)
    # below the threshold: nothing added
    g_dim = gbuffer(W, H, normal=(0.0, 0.0, 1.0), rough=0.4, metal=0.3, em=(0.01, 0.01, 0.01))
    np.testing.assert_array_equal(lit(W, H, g_dim, (0.2, 0.3, -0.9)), lit(W, H, g, (0.2, 0.3, -0.9)))


def test_background_passes_colour():
    W, H = 3, 2
    rgba, dep, nw, mat = gbuffer(W, H, color=(0.3, 0.2, 0.1, 0.7))
    dep[0, 0], dep[0, 1], dep[1, 2] = 1.0, 0.0, 0.99995
    out = half_bits_to_f32(lit(W, H, (rgba, dep, nw, mat), (0.0, -1.0, 0.0)))
    for y, x in ((0, 0), (0, 1), (1, 2)):
        np.testing.assert_array_equal(out[y, x], [f16(0.3), f16(0.2), f16(0.1), 0.0])
    assert out[1, 0, 3] == f16(0.7)


def lift_from(out, c_alb, ambient=0.2):
    """Invert the ambient-only output srgb(reinhard(ambient * albedo * ao)) for ao (green channel)."""
    s = out[..., 1].astype(np.float64) ** 2.2
    return (s / (1.0 - s)) / (ambient * c_alb)


def test_bloom_sample_lifts_ao():
    W, H = 8, 6
    rgba, dep, nw, mat = gbuffer(W, H, color=(0.5, 0.5, 0.5, 1.0))
    c = f16(0.5)
    lum = np.float64(c)  # grey: luminance weights sum to 1
    scale = lut_of(1.0, 0.0, 1.0)[int(np.float32(lum) * 255)] / lum
    alb_g = (c * scale) ** 2.2
    ao0 = np.zeros((H, W, 4), np.float32)  # AO fully occluded: only the lift remains
    bw, bh = (W + 1) // 2, (H + 1) // 2
    # constant field v: lift = 2 v, the same as an AO texture of 2 v
    v = 0.2
    bloom = np.full((bh, bw, 4), v, np.float32)
    a = half_bits_to_f32(ambient_only(W, H, rgba, dep, nw, mat, ao=ao0, bloom=bloom))
    aotex = np.full((H, W, 4), 2 * v, np.float32)
    b = half_bits_to_f32(ambient_only(W, H, rgba, dep, nw, mat, ao=aotex))
    np.testing.assert_allclose(a, b, atol=1e-3)
    # a field linear in the column, B(i) = 0.05 + 0.04 i (grey): the sample at texel coordinate
    # tx = (x + 0.5) / 2 - 0.5 is B(tx) inside; column 0 (tx = -0.25) wraps to the last column
    # with weight 1/4; rows clamp (the field does not depend on them)
    i = np.arange(bw, dtype=np.float32)
    field = (0.05 + 0.04 * i).astype(np.float32)
    bloom = np.broadcast_to(field[None, :, None], (bh, bw, 4)).copy()
    out = half_bits_to_f32(ambient_only(W, H, rgba, dep, nw, mat, ao=ao0, bloom=bloom))
    lift = lift_from(out, alb_g)
    for x in range(W):
        tx = (x + 0.5) / 2 - 0.5
        if x == 0:
            want = 0.25 * field[-1] + 0.75 * field[0]
        elif tx > bw - 1:
            want = (1 - (tx - (bw - 1))) * field[-1] + (tx - (bw - 1)) * field[0]  # wraps right
        else:
            want = 0.05 + 0.04 * tx
        np.testing.assert_allclose(lift[:, x], 2 * want, rtol=4e-3)
