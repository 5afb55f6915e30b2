"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs of the C++ that parses untrusted input
(the PLY decoder, csrc/assets.cpp) and of the CPU oracle (oracle/hdsr_oracle.cpp).

Both are built by tests/sanitize/Makefile with g++ -fsanitize=address,undefined
-fno-sanitize-recover=all into small drivers that copy every input into a heap buffer of exactly
its size, so any read or write past a buffer, any use after free and any undefined arithmetic
stops the process with a report.
  * PLY: valid files (all-float 3DGS rows with SH, normals and material fields; byte colours; every
    scalar type) decode with the expected counts, and a seeded malformed corpus (truncations at
    every offset band, bad counts, formats, property lines, NUL bytes, huge headers, random header
    byte flips) is rejected with an error code - the decoder's messages are the reference's
    (src/localPly.ts:488-598, 640-721).
  * Oracle: seeded frames in the plain, G-buffer, fp64, differentiable and f16 modes, at a
    viewport that is not a multiple of the tile, at a one-pixel viewport and with a dense core
    that makes tiles exceed 4096 entries; every output is copied out.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from hdsr import camera
from scenes_for_tests import dense_scene, frame_inputs

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")


@pytest.fixture(scope="module")
def drivers():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ / make not available")
    r = subprocess.run(["make", "-s", "-C", SAN, "-j4"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return os.path.join(SAN, "build", "ply_driver"), os.path.join(SAN, "build", "oracle_driver")


def run_clean(cmd, timeout=600):
    r = subprocess.run(cmd, capture_output=True, encoding="utf-8", errors="replace", timeout=timeout, env=ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    return r.stdout


CODES = {"float": "<f4", "double": "<f8", "uchar": "u1", "char": "i1", "short": "<i2", "ushort": "<u2",
         "int": "<i4", "uint": "<u4"}
BASE = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
        "rot_0", "rot_1", "rot_2", "rot_3"]


def ply(props, rows, fmt="binary_little_endian 1.0", count=None, extra=()):
    head = ["ply", f"format {fmt}", *extra, f"element vertex {len(rows) if count is None else count}"]
    head += [f"property {t} {n}" for t, n in props]
    head += ["end_header", ""]
    data = np.array([tuple(r) for r in rows], dtype=np.dtype([(n, CODES[t]) for t, n in props]))
    return "\n".join(head).encode() + data.tobytes()


def valid_files(rng):
    n = 37
    names = BASE + [f"f_rest_{i}" for i in range(45)] + ["nx", "ny", "nz", "roughness", "metallic",
                                                         "emissive_r", "emissive_g", "emissive_b",
                                                         "detail_normal_x", "detail_normal_y", "detail_normal_z"]
    full = ply([("float", k) for k in names], rng.normal(0, 1, (n, len(names))).tolist(),
               extra=("comment made by a seeded generator", "obj_info test"))
    plain = ply([("float", k) for k in BASE], rng.normal(0, 1, (5, len(BASE))).tolist())
    no_dc = [k for k in BASE if not k.startswith("f_dc")]
    byte_col = ply([("float", k) for k in no_dc] + [("uchar", "red"), ("uchar", "green"), ("uchar", "blue")],
                   [tuple(rng.normal(0, 1, len(no_dc))) + (10, 200, 255), tuple(rng.normal(0, 1, len(no_dc))) + (0, 0, 0)])
    types = [(t, f"p{i}") for i, t in enumerate(CODES)]
    every = ply([(t, k) for t, k in zip(list(CODES) * 2, BASE)] + types,
                [tuple(range(1, len(BASE) + 1)) + tuple(range(1, 9)) for _ in range(3)])
    many = ply([("float", k) for k in BASE], rng.normal(0, 1, (150, len(BASE))).tolist())  # percentile bounds
    return {"full.ply": (full, n), "plain.ply": (plain, 5), "bytecol.ply": (byte_col, 2), "types.ply": (every, 3),
            "many.ply": (many, 150)}


def malformed_files(rng, good):
    out = {"empty.ply": b"", "one.ply": b"p", "magic.ply": b"ply", "magic_nl.ply": b"ply\n",
           "noend.ply": good[:good.index(b"end_header")]}
    cut = good.index(b"end_header\n") + len(b"end_header\n")
    for k in list(range(0, cut, 7)) + [cut, cut + 1, cut + 13, len(good) - 5, len(good) - 1]:
        out[f"trunc_{k}.ply"] = good[:k]
    for i, bad in enumerate((b"4294967295", b"4294967296", b"99999999999999999999999", b"-5", b"abc", b"1e3", b"",
                             b"2.5", b"Infinity", b"NaN")):
        out[f"count_{i}.ply"] = good.replace(b"element vertex 37", b"element vertex " + bad)
    # counts JavaScript's Number() accepts (src/localPly.ts:508): a hex literal, extra spaces
    out["count_hex.ply"] = good.replace(b"element vertex 37", b"element vertex 0x10")
    out["count_space.ply"] = good.replace(b"element vertex 37", b"element vertex  7")
    for i, fmt in enumerate((b"ascii 1.0", b"binary_big_endian 1.0", b"binary_little_endian 2.0", b"")):
        out[f"fmt_{i}.ply"] = good.replace(b"binary_little_endian 1.0", fmt)
    out["fmt_bare.ply"] = good.replace(b"binary_little_endian 1.0", b"binary_little_endian")  # accepted (:534)
    out["face_first.ply"] = good.replace(b"element vertex", b"element face 3\nproperty list uchar int i\nelement vertex")
    out["face_after.ply"] = good.replace(b"end_header", b"element face 3\nproperty list uchar int vertex_indices\n"
                                                        b"end_header")
    out["list_prop.ply"] = good.replace(b"property float x", b"property list uchar float x")
    out["short_prop.ply"] = good.replace(b"property float y", b"property float")
    out["bad_type.ply"] = good.replace(b"property float z", b"property float64x z")
    out["dup_prop.ply"] = good.replace(b"property float y", b"property float x")
    out["no_props.ply"] = b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nend_header\n" + b"\0" * 64
    out["nul_header.ply"] = good[:40] + b"\0" + good[41:]
    out["long_line.ply"] = good.replace(b"comment made", b"comment " + b" " * 70000 + b"made")
    out["many_props.ply"] = (b"ply\nformat binary_little_endian 1.0\nelement vertex 4294967295\n"
                             + b"".join(b"property double q%d\n" % i for i in range(3000))
                             + b"property float x\nproperty float y\nproperty float z\nend_header\n" + b"\0" * 256)
    out["f_rest_47.ply"] = ply([("float", k) for k in ["x", "y", "z"] + [f"f_rest_{i}" for i in range(47)]],
                               [(0.0,) * 50])
    out["f_rest_gap.ply"] = ply([("float", k) for k in ["x", "y", "z", "f_rest_0", "f_rest_1", "f_rest_5"]],
                                [(0.0,) * 6])
    out["crlf_cut.ply"] = good[:cut].replace(b"\n", b"\r\n")
    head = bytearray(good[:cut])
    for i in range(200):
        m = bytearray(head)
        for _ in range(int(rng.integers(1, 4))):
            m[int(rng.integers(0, len(m)))] = int(rng.integers(0, 256))
        out[f"flip_{i}.ply"] = bytes(m) + good[cut:]
    return out


def test_ply_decoder_under_sanitizers(drivers, tmp_path):
    ply_driver, _ = drivers
    rng = np.random.default_rng(5)
    good = valid_files(rng)
    bad = malformed_files(rng, good["full.ply"][0])
    paths = []
    for name, data in list((k, v[0]) for k, v in good.items()) + list(bad.items()):
        p = tmp_path / name
        p.write_bytes(data)
        paths.append(str(p))
    lines = run_clean([ply_driver, *paths]).strip().splitlines()
    assert len(lines) == len(paths)
    res = {os.path.basename(p): ln.split(" ", 2) for p, ln in zip(paths, lines)}
    for name, (_, count) in good.items():
        assert res[name][0] == "0" and int(res[name][1]) == count, (name, res[name])
    assert res["count_hex.ply"][:2] == ["0", "16"] and res["count_space.ply"][:2] == ["0", "7"]
    # accepted as the reference's header parse accepts them (src/localPly.ts:501-536): elements
    # other than vertex are skipped, a bare format name passes, NUL / long comment lines are text
    accepted = ("long_line.ply", "count_hex.ply", "count_space.ply", "fmt_bare.ply", "face_first.ply",
                "face_after.ply", "nul_header.ply")
    for name in accepted:
        assert res[name][0] == "0", (name, res[name])
    must_fail = [k for k in bad if not k.startswith("flip_") and k not in accepted]
    for name in must_fail:
        assert int(res[name][0]) < 0, (name, res[name])
        assert res[name][2].strip(), name  # a message accompanies every error
    assert sum(int(res[k][0]) < 0 for k in bad if k.startswith("flip_")) > 50
    print(f"{len(paths)} files: {sum(int(v[0]) < 0 for v in res.values())} rejected, none tripped a sanitizer")


def write_inputs(d, scene, srt, u, sh_degree, logits=False):
    os.makedirs(d, exist_ok=True)
    op = scene.opacities.astype(np.float64)
    opac = np.log(op / (1.0 - op)).astype(np.float32) if logits else scene.opacities
    arrays = {"positions": scene.positions, "scales": scene.scales, "rotations": scene.rotations,
              "opacities": opac, "sorted": srt, "materials": scene.materials, "shdata": scene.sh_data(),
              "uniforms": u}
    for k, v in arrays.items():
        np.ascontiguousarray(v).tofile(os.path.join(d, k + ".bin"))


@pytest.mark.parametrize("width,height,flags", [(200, 136, (0, 1, 2, 32)), (200, 136, (16,)), (1, 1, (0, 1, 16))])
def test_oracle_under_sanitizers(drivers, tmp_path, width, height, flags):
    _, oracle_driver = drivers
    scene = dense_scene(n=9000, seed=21, sh_degree=3)
    cam = camera.Camera(camera.orbit_eye(2.0, 0.4, 0.2), (0, 0, 0), 180.0, 180.0, width, height)
    plan, srt, u = frame_inputs(scene, cam)
    d = str(tmp_path / "in")
    write_inputs(d, scene, srt, u, 3, logits=(flags == (16,)))
    out = run_clean([oracle_driver, d, str(width), str(height), "3", *map(str, flags)])
    rows = [ln.split() for ln in out.strip().splitlines()]
    assert [int(r[0]) for r in rows] == list(flags)
    if width > 1:
        assert all(int(r[1]) > 9000 for r in rows)  # the dense core overlaps many tiles
