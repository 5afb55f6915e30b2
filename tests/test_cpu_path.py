"""The JavaScript restatement of the reference's CPU path (oracle/js/cpu_path.js, timed as the CPU
baseline) renders the same image as the fp32 oracle up to double-vs-float rounding."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


def write_inputs(d, s, u, W, H, write_rgba=True):
    for k, v in {"positions.bin": s.positions, "scales.bin": s.scales, "rotations.bin": s.rotations,
                 "opacities.bin": s.opacities, "shdata.bin": s.sh_data(), "uniforms.bin": u}.items():
        np.ascontiguousarray(v).tofile(os.path.join(d, k))
    json.dump({"count": s.count, "width": W, "height": H, "shDegree": s.sh_degree, "writeRgba": write_rgba},
              open(os.path.join(d, "meta.json"), "w"))


@pytest.mark.skipif(NODE is None, reason="node missing")
@pytest.mark.parametrize("sh3", [False, True])
def test_js_cpu_path_matches_oracle(sh3):
    s = dense_scene(n=6000, seed=31, sh_degree=3 if sh3 else 0, core_fraction=0.5)
    cam = camera.Camera(camera.orbit_eye(3.0, 0.3, 0.2), (0, 0, 0), 525.0, 525.0, 320, 256)
    plan = compositor.planTileSplatCompositor({"viewportWidth": 320, "viewportHeight": 256, "tileSizePx": 16,
                                               "splatCount": s.count})
    u = camera.frame_uniforms_for(cam, plan, s.sh_degree)
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    orc = oracle.run(s, srt, u, 320, 256)
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, s, u, 320, 256)
        r = subprocess.run([NODE, os.path.join(ROOT, "oracle", "js", "cpu_path.js"), d, "1"], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        t = json.loads(r.stdout.strip().splitlines()[-1])
        rgba = np.fromfile(os.path.join(d, "rgba_cpu.bin"), np.float32)
    assert t["refs"] == orc.stats["refs"]
    diff = np.abs(rgba - orc["rgba"])
    # double vs fp32 arithmetic: tiny everywhere except rare threshold flips
    assert np.median(diff) < 1e-6
    assert np.mean(diff > 1e-3) < 0.01
    for k in ("cpuDepthSort_s", "project_s", "tile_refs_s", "composite_sample_s"):
        assert t[k] >= 0
