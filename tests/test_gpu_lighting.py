"""Deferred lighting on the GPU (hdsr_deferred_lighting) against the oracle's restatement, bit for
bit on the rgba16float output.

A 60k-splat SH3 clusters-and-sheets scene with random roughness / metalness and emissive on a tenth
of the splats is rendered at 640x360 with the G-buffer; the normal passes run first, as the
reference encodes them (src/splatRenderer.ts:1042-1069).  The oracle gets the frame's own colour,
depth, post-filter normal words and material words, the library's Params block and curve tables.
Cases: the reference viewer's defaults (no AO texture, no bloom, identity curves); a random AO
texture, a random half-resolution bloom texture, non-identity curves, another exposure; the
specular-only mode; a separate lighting camera (lightingViewProj / lightingCameraPosition).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from hdsr import _lib, camera, compositor, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

W, H = 640, 360


@pytest.fixture(scope="module")
def frame(hip_device):
    s = scenes.scene_config1(60_000, seed=41)
    rng = np.random.default_rng(42)
    s.materials = scenes.pack_material(s.count, rng.uniform(0, 1, s.count), rng.uniform(0, 1, s.count))
    em = np.zeros((s.count, 3), np.float32)
    pick = rng.random(s.count) < 0.1
    em[pick] = rng.uniform(0, 1, (int(pick.sum()), 3))
    s.emissive = em
    cam = camera.Camera(camera.orbit_eye(3.0, 0.5, 0.25), (0, 0, 0), 420.0, 420.0, W, H)
    plan = compositor.planTileSplatCompositor({"viewportWidth": W, "viewportHeight": H, "tileSizePx": 16,
                                               "splatCount": s.count})
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    res = compositor.createTileSplatCompositor(None, plan, gbuffer=True)
    compositor.bind_scene(res, s, srt)
    res.set_frame(u)
    compositor.encodeFullComputeCompositorPipeline(None, res)
    res.encodeScreenSpaceNormals(None)
    res.encodeBilateralNormalFilter(None)
    g = {"rgba": res.read_output(_lib.OUT_RGBA, np.float32, (H, W, 4)),
         "depth": res.read_depth(),
         "normal": res.read_output(_lib.OUT_GB_NORMAL, np.uint32, (H, W)),
         "mat": res.read_output(_lib.OUT_GB_MATERIAL, np.uint32, (H, W, 4))}
    covered = g["depth"] < 0.9999
    assert covered.mean() > 0.2
    assert np.count_nonzero(g["mat"][..., 1][covered]) > 100  # emissive reaches the G-buffer
    yield res, u, g, hip_device
    res.destroy()


def host_block(p, vp16, cam3):
    inv, ok = oracle.mat4_inverse(vp16)
    assert ok
    out = np.zeros(44, np.float32)
    c = np.ascontiguousarray(cam3, np.float32)
    _lib.load().hdsr_write_lighting_params(C.byref(p), inv.ctypes.data, c.ctypes.data, W, H, out.ctypes.data)
    return out


def luts(p):
    out = np.zeros(768, np.float32)
    lib = _lib.load()
    for i, cv in enumerate((p.roughness_curve, p.metalness_curve, p.albedo_curve)):
        lib.hdsr_material_curve_lut(C.byref(cv), out[i * 256:].ctypes.data)
    return out


def run_and_check(res, u, g, p, ao=None, bloom=None, vp16=None, cam3=None):
    res.encodeDeferredLighting(None, p)
    got = res.read_output(_lib.OUT_LIT, np.uint16, (H, W, 4))
    P = host_block(p, u[:16] if vp16 is None else vp16, u[28:31] if cam3 is None else cam3)
    want = oracle.deferred_lighting(g["rgba"], g["depth"], g["normal"], g["mat"], P, luts(p), ao=ao, bloom=bloom)
    bad = np.argwhere(got != want)
    assert len(bad) == 0, (len(bad), bad[:5], got[tuple(bad[0])] if len(bad) else None)
    return got


def test_lighting_defaults_match_oracle(frame):
    res, u, g, dev = frame
    p = res.lighting_params()
    got = run_and_check(res, u, g, p)
    o = _lib.Outputs()
    assert _lib.load().hdsr_get_outputs(res.handle, C.byref(o)) == 0 and o.lit
    lit = got.view(np.float16).astype(np.float32)
    covered = g["depth"] < 0.9999
    assert np.all(np.isfinite(lit)) and lit[covered][:, :3].max() > 0.05
    # background pixels pass the colour through with alpha 0
    bg = ~covered & (g["depth"] > 0.0001)
    if bg.any():
        assert np.all(lit[bg][:, 3] == 0.0)


def test_lighting_with_ao_bloom_curves(frame):
    res, u, g, dev = frame
    rng = np.random.default_rng(7)
    ao = rng.uniform(0, 1, (H, W, 4)).astype(np.float16).astype(np.float32)
    bloom = rng.uniform(0, 0.4, ((H + 1) // 2, (W + 1) // 2, 4)).astype(np.float16).astype(np.float32)
    ao_d = torch.from_numpy(ao).to(dev)
    bloom_d = torch.from_numpy(bloom).to(dev)
    p = res.lighting_params(light_direction=(-0.4, -0.8, 0.3), light_intensity=4.5, ambient_intensity=0.3,
                            exposure=1.7, emissive_intensity=2.0, emissive_threshold=0.3,
                            roughness_curve=(1.4, -0.1, 0.8), metalness_curve=(0.7, 0.2, 1.3),
                            albedo_curve=(1.2, 0.05, 1.1), ao=ao_d.data_ptr(), bloom=bloom_d.data_ptr())
    run_and_check(res, u, g, p, ao=ao, bloom=bloom)
    p.specular_only = 1
    run_and_check(res, u, g, p, ao=ao, bloom=bloom)
    torch.cuda.synchronize()


def test_lighting_separate_lighting_camera(frame):
    res, u, g, dev = frame
    other = camera.Camera(camera.orbit_eye(3.5, 1.2, 0.4), (0, 0, 0), 420.0, 420.0, W, H)
    plan = compositor.planTileSplatCompositor({"viewportWidth": W, "viewportHeight": H, "tileSizePx": 16,
                                               "splatCount": 60_000})
    u2 = camera.frame_uniforms_for(other, plan, 3)
    p = res.lighting_params(lighting_view_proj=u2[:16], lighting_camera_position=u2[28:31])
    run_and_check(res, u, g, p, vp16=u2[:16], cam3=u2[28:31])
