"""Host-side checks that need no GPU: the C ABI library loads and exports every symbol the header
declares, the plan arithmetic, and the reference's own known-answer tests for the host helpers
(cpuDepthSort orders, the viewer flip, buffer layouts, material packing, SH coefficient layout)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from hdsr import _lib, camera, compositor, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(hdsr_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    _lib.load()  # through the binding first: it puts torch's HIP runtime in place
    lib = C.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), f"missing export {n}"
    declared = {n for n, _, _ in _lib.SIGNATURES}
    assert declared == set(names), f"binding table differs from header: {declared ^ set(names)}"
    assert _lib.load().hdsr_version().startswith(b"hdsr")


def test_one_hip_runtime_when_loaded_before_torch():
    """The binding loaded before anything imports torch still leaves one libamdhip64 mapped: torch
    loads its own copy by path, and of two HIP runtimes in one process the one that initialises
    second finds no device."""
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(_lib.__file__)))
    code = ("import sys\nsys.path.insert(0, %r)\nfrom hdsr import _lib\n_lib.load()\nimport torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(sorted({l.split()[-1] for l in maps if 'libamdhip64' in l}))" % pkg)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert len(eval(r.stdout.strip().splitlines()[-1])) == 1, r.stdout


def test_plan_matches_reference_arithmetic():
    # planTileSplatCompositor (gpuTileSplatCompositor.ts:109-139)
    cases = [((640, 480, 50_000), (40, 30, 1200, 1720, 800_000)),
             ((1920, 1080, 1_000_000), (120, 68, 8160, 13600, 16_000_000)),
             ((3840, 2160, 6_000_000), (240, 135, 32400, 54398, 65535 * 256)),
             ((1280, 720, 500_000), (80, 45, 3600, 6390, 8_000_000))]
    for (w, h, n), exp in cases:
        p = compositor.planTileSplatCompositor({"viewportWidth": w, "viewportHeight": h, "tileSizePx": 16,
                                                "splatCount": n})
        assert (p["tileColumns"], p["tileRows"], p["tileCount"], p["mortonTileCount"], p["maxTotalTileRefs"]) == exp
    p = compositor.planTileSplatCompositor({"viewportWidth": 100, "viewportHeight": 50, "tileSizePx": 16,
                                            "splatCount": 10, "tileEntryMultiplier": 2.5})
    assert p["maxTotalTileRefs"] == 25


def test_plan_rejects_bad_input():
    with pytest.raises(_lib.HdsrError):
        compositor.planTileSplatCompositor({"viewportWidth": 0, "viewportHeight": 10, "tileSizePx": 16,
                                            "splatCount": 1})


def test_create_rejects_unsupported_options_without_gpu_work():
    p = compositor.planTileSplatCompositor({"viewportWidth": 64, "viewportHeight": 64, "tileSizePx": 16,
                                            "splatCount": 1})
    h = C.c_void_p()
    for flags in (_lib.HDSR_DIFFERENTIABLE | _lib.HDSR_REF_CAPS, _lib.HDSR_DIFFERENTIABLE | _lib.HDSR_F16_COMPOSITE,
                  _lib.HDSR_DIFFERENTIABLE | _lib.HDSR_GBUFFER):
        rc = _lib.load().hdsr_create(C.byref(compositor._plan_struct(p)), flags, C.byref(h))
        assert rc == _lib.HDSR_E_UNSUPPORTED and not h.value
    bad = dict(p, tileSizePx=8)
    rc = _lib.load().hdsr_create(C.byref(compositor._plan_struct(bad)), 0, C.byref(h))
    assert rc == _lib.HDSR_E_INVALID_ARG
    # more than 255 tiles on an axis does not fit the u8 tile-box packing (gpu_project_splats.wgsl:229-231)
    for w, hgt in ((4096, 64), (64, 4096)):
        wide = compositor.planTileSplatCompositor({"viewportWidth": w, "viewportHeight": hgt, "tileSizePx": 16,
                                                   "splatCount": 1})
        rc = _lib.load().hdsr_create(C.byref(compositor._plan_struct(wide)), 0, C.byref(h))
        assert rc == _lib.HDSR_E_INVALID_ARG and not h.value
    edge = compositor.planTileSplatCompositor({"viewportWidth": 4080, "viewportHeight": 16, "tileSizePx": 16,
                                               "splatCount": 1})
    assert edge["tileColumns"] == 255


IDENT = np.eye(4, dtype=np.float32).T.reshape(16)


def test_cpu_depth_sort_kats():
    # tests/cpuDepthSort.test.ts:18-84
    pos = np.array([[0, 0, -2], [0, 0, -9], [0, 0, -5], [0, 0, -0.5]], np.float32)
    assert scenes.sort_back_to_front(pos, IDENT).tolist() == [1, 2, 0, 3]
    trans = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, -5, 1], np.float32)
    pos = np.array([[0, 0, 0], [3, 2, 2], [-1, 1, -4]], np.float32)
    assert scenes.sort_back_to_front(pos, trans).tolist() == [2, 0, 1]
    xdepth = np.array([1, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32)
    pos = np.array([[-4, 0, 0], [-1, 0, 0], [-6, 0, 0]], np.float32)
    assert scenes.sort_back_to_front(pos, xdepth).tolist() == [2, 0, 1]
    pos = np.array([[0, 0, -4], [0, 0, -6], [1, 0, -4], [2, 0, -4]], np.float32)
    assert scenes.sort_back_to_front(pos, IDENT).tolist() == [1, 0, 2, 3]
    with pytest.raises(ValueError):
        scenes.sort_back_to_front(np.array([[np.nan, 0, 0]], np.float32), IDENT)


def test_viewer_flip_kat():
    # tests/renderer/realSmokeScene.test.mjs:149-163
    vp = camera.view_projection(IDENT, IDENT)
    assert vp.tolist() == [1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]


def test_camera_conventions():
    cam = camera.Camera((0, 0, 3), (0, 0, 0), 525, 525, 640, 480)
    V = camera.col_major(cam.view)
    assert np.allclose(V @ [0, 0, 0, 1], [0, 0, -3, 1])       # looks down -z
    assert np.allclose(V @ [1, 0, 0, 1], [1, 0, -3, 1])       # +x right
    P = camera.col_major(cam.proj)
    clip = P @ [0, 0, -camera.NEAR, 1]
    assert abs(clip[2] / clip[3]) < 1e-6                      # near plane -> depth 0
    clip = P @ [0, 0, -camera.FAR, 1]
    assert abs(clip[2] / clip[3] - 1) < 1e-6                  # far plane -> depth 1
    plan = compositor.planTileSplatCompositor({"viewportWidth": 640, "viewportHeight": 480, "tileSizePx": 16,
                                               "splatCount": 7})
    u = camera.frame_uniforms_for(cam, plan, 3)
    assert u.nbytes == compositor.TILE_SPLAT_FRAME_UNIFORM_BYTES == 256
    w = u.view(np.uint32)
    assert (u[16], u[17], u[18]) == (640, 480, 16)
    assert (w[20], w[21], w[22], w[23], w[25]) == (40, 30, 7, plan["maxTotalTileRefs"], 3)
    assert (u[48], u[49]) == (525.0, 525.0)                  # focal = W * P[0] / 2
    assert u[52] == 1 and u[57] == 1 and u[62] == 1 and u[55] == 0  # identity normalMatrix, vec4 columns
    np.testing.assert_array_equal(u[28:31], np.float32([0, 0, 3]))


def test_material_and_sh_packing_kats():
    # splats.test.ts:375-431: default material pack2x16float(0.75, 0), shData stride 3 + 3K + 3
    m = scenes.pack_material(3)
    assert np.all(m == 0x3A00)
    # truncating f32ToF16 flushes below 2^-14 and truncates the mantissa (splats.ts:431-442)
    assert scenes.f32_to_f16_trunc(np.float32([1e-5]))[0] == 0
    assert scenes.f32_to_f16_trunc(np.float32([1.0009765]))[0] == 0x3C00
    assert scenes.f32_to_f16_trunc(np.float32([70000.0]))[0] == 0x7C00
    colors = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], np.float32)
    sd = scenes.pack_sh_data(colors)
    assert sd.tolist() == pytest.approx([0.1, 0.2, 0.3, 0, 0, 0, 0.4, 0.5, 0.6, 0, 0, 0])
    sh = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    em = np.ones((2, 3), np.float32)
    sd = scenes.pack_sh_data(colors, sh, em).reshape(2, -1)
    assert sd.shape[1] == 3 + 9 + 3
    np.testing.assert_array_equal(sd[1, 3:12], sh[1].reshape(-1))
    np.testing.assert_array_equal(sd[:, 12:], em)


def test_ply_f_rest_to_coefficient_major_kat():
    # preprocessing/tests/test_sh_payload_export.py:15-33: PLY f_rest is channel-major
    f_rest = np.array([[10, 11, 12, 20, 21, 22, 30, 31, 32]], np.float32)
    sh = scenes.sh_from_ply_f_rest(f_rest)
    np.testing.assert_array_equal(sh, np.array([[[10, 20, 30], [11, 21, 31], [12, 22, 32]]], np.float32))


def test_gpu_sort_order_restatement():
    # the vectorised network equals the reference's loop (restated pair by pair) on inputs with
    # ties, signed zeros, infinities and non-power-of-two sizes
    from oracle import oracle as orc
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 7, 16, 33, 100):
        for keys in (rng.choice(np.array([-2.5, -1, -0.0, 0.0, 0.5, 3, np.inf, -np.inf], np.float32), n),
                     rng.normal(0, 1, n).astype(np.float32)):
            assert orc.gpu_sort_order(keys).tolist() == orc.bitonic_sort_restated(keys).tolist()
    # distinct keys: a plain descending sort; the cpuDepthSort known answers hold
    keys = rng.permutation(1000).astype(np.float32)
    assert orc.gpu_sort_order(keys).tolist() == np.argsort(-keys, kind="stable").tolist()
    for pos, want in [(np.array([[0, 0, -2], [0, 0, -9], [0, 0, -5], [0, 0, -0.5]], np.float32), [1, 2, 0, 3]),
                      (np.array([[0, 0, -4], [0, 0, -6], [1, 0, -4], [2, 0, -4]], np.float32), [1, 0, 2, 3])]:
        assert orc.gpu_sort_order(orc.view_depth_keys_f32(pos, IDENT)).tolist() == want
    # the network's tie order is not "ties to the lower index": equal keys can come out of order
    keys = np.array([np.inf, 3, -np.inf, -2.5, -2.5, np.inf, 0, 3, 0, -np.inf, -2.5, 3, -2.5, -1, np.inf, -0.0],
                    np.float32)
    assert orc.gpu_sort_order(keys).tolist()[6:9] == [6, 15, 8]
    # against the float64 CPU sort: identical except around f32 ties
    sc = scenes.scene_config0(20000, seed=4)
    cam = scenes.cameras_config0()[0]
    a = orc.gpu_sort_order(orc.view_depth_keys_f32(sc.positions, cam.view))
    b = scenes.sort_back_to_front(sc.positions, cam.view)
    assert np.mean(a != b) < 0.01


# FrameUniforms (src/shaders/gpu_project_splats.wgsl:27-44): WGSL field, byte offset, word type,
# words.  mat3x3f columns are vec3 with a 16-byte stride (a pad word after each column).
FRAME_UNIFORMS_LAYOUT = [
    ("viewProj", 0, "f32", 16), ("viewport", 64, "f32", 2), ("tileSizePx", 72, "f32", 1),
    ("debugMode", 76, "f32", 1), ("tileGrid", 80, "u32", 2), ("splatCount", 88, "u32", 1),
    ("totalTileRefs", 92, "u32", 1), ("splatScale", 96, "f32", 1), ("shDegree", 100, "u32", 1),
    ("_pad0", 104, "u32", 2), ("cameraPos", 112, "f32", 3), ("_pad1", 124, "f32", 1),
    ("viewMatrix", 128, "f32", 16), ("focal", 192, "f32", 2), ("_pad2", 200, "f32", 2),
    ("normalMatrix.c0", 208, "f32", 3), ("_padc0", 220, "f32", 1), ("normalMatrix.c1", 224, "f32", 3),
    ("_padc1", 236, "f32", 1), ("normalMatrix.c2", 240, "f32", 3), ("_padc2", 252, "f32", 1),
]


def test_frame_uniform_layout_word_by_word():
    """The 256-byte block against the WGSL struct's offsets, field by field, with the values
    writeTileSplatFrameUniforms (src/gpuTileSplatCompositor.ts:730-774) stores there: viewport and
    tile size as f32, the grid / counts / degree as u32, focal = (W * P[0] * 0.5, H * P[5] * 0.5)
    rounded once to f32, normalMatrix columns at a 16-byte stride, every pad word zero.  The JS
    writer equals this one word for word (test_node.py::test_facade_cpu)."""
    words = [w for _, _, _, n in FRAME_UNIFORMS_LAYOUT for w in range(n)]
    assert len(words) == 64
    off = 0
    for name, byte, _, n in FRAME_UNIFORMS_LAYOUT:  # contiguous, no overlap, 256 bytes
        assert byte == off, name
        off += 4 * n
    assert off == 256
    plan = compositor.planTileSplatCompositor({"viewportWidth": 1000, "viewportHeight": 563, "tileSizePx": 16,
                                               "splatCount": 123457})
    vp = np.float32(np.arange(16) * 0.25 - 1.5)
    view = np.float32(np.arange(16) * -0.125 + 3.0)
    proj = np.zeros(16, np.float32)
    proj[0], proj[5] = np.float32(1.7320508), np.float32(2.4142137)
    nm = [0.5, -0.25, 0.125, 2.0, 3.0, -4.0, 7.0, 8.0, 9.5]
    t = camera.write_frame_uniforms(vp, plan, 1.25, 2, [1.5, -2.5, 3.5], view, proj, nm)
    assert t.dtype == np.float32 and t.nbytes == 256
    want = {
        "viewProj": vp, "viewport": [1000.0, 563.0], "tileSizePx": [16.0], "debugMode": [0.0],
        "tileGrid": [63, 36], "splatCount": [123457], "totalTileRefs": [min(16 * 123457, 65535 * 256)],
        "splatScale": [1.25], "shDegree": [2], "_pad0": [0, 0], "cameraPos": [1.5, -2.5, 3.5], "_pad1": [0.0],
        "viewMatrix": view, "focal": [1000 * float(proj[0]) * 0.5, 563 * float(proj[5]) * 0.5], "_pad2": [0.0, 0.0],
        "normalMatrix.c0": nm[0:3], "_padc0": [0.0], "normalMatrix.c1": nm[3:6], "_padc1": [0.0],
        "normalMatrix.c2": nm[6:9], "_padc2": [0.0],
    }
    u = t.view(np.uint32)
    for name, byte, kind, n in FRAME_UNIFORMS_LAYOUT:
        w = byte // 4
        exp = np.asarray(want[name], np.uint32 if kind == "u32" else np.float32)
        got = u[w:w + n] if kind == "u32" else t[w:w + n]
        np.testing.assert_array_equal(got, exp, err_msg=name)
    assert plan["tileColumns"] == 63 and plan["tileRows"] == 36


class _FakeResources:
    """Stands in for a context: the facade's length checks run before any call into the library."""
    plan = {"splatCount": 10}
    handle = None


@pytest.mark.parametrize("key,per", [("positionBuffer", 3), ("scaleBuffer", 3), ("rotationBuffer", 4),
                                     ("opacityBuffer", 1), ("materialBuffer", 1), ("sortedIndexBuffer", 1)])
def test_bind_rejects_short_buffers(key, per):
    n = 10
    bufs = {"positionBuffer": np.zeros(3 * n, np.float32), "scaleBuffer": np.zeros(3 * n, np.float32),
            "rotationBuffer": np.zeros(4 * n, np.float32), "opacityBuffer": np.zeros(n, np.float32),
            "materialBuffer": np.zeros(n, np.uint32), "shDataBuffer": np.zeros(6 * n, np.float32),
            "sortedIndexBuffer": np.arange(n, dtype=np.uint32)}
    bufs[key] = bufs[key][:-1]
    with pytest.raises(ValueError, match=key):
        compositor.createTileSplatBindGroups(None, _FakeResources(), bufs)
