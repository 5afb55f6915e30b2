"""Asset decoding (host only): the C++ PLY decoder and the first-smoke manifest decoder.

Known answers from the reference's own tests (tests/localPly.test.ts:8-178), a PLY round trip
against an independent numpy decode (SH reorder by hdsr.scenes.sh_from_ply_f_rest), the error
messages of src/localPly.ts and src/splats.ts, and the committed subset of the reference's real
first-smoke scene (tests/golden/smoke_subset, made by tests/golden/make_smoke_subset.py).
"""
import json
import math
import os
import shutil

import numpy as np
import pytest

from hdsr import assets

SMOKE = os.path.join(os.path.dirname(__file__), "golden", "smoke_subset", "scaniverse-first-smoke.json")
SH_C0_TEST = 0.28209479  # the constant tests/localPly.test.ts:6 uses


def ply_bytes(props, rows, fmt="binary_little_endian 1.0", extra_header=()):
    """A PLY with one vertex element; props = [(type, name)], rows = list of value tuples."""
    codes = {"float": "<f4", "double": "<f8", "uchar": "u1", "char": "i1", "short": "<i2", "ushort": "<u2",
             "int": "<i4", "uint": "<u4"}
    head = ["ply", f"format {fmt}", *extra_header, f"element vertex {len(rows)}"]
    head += [f"property {t} {n}" for t, n in props]
    head += ["end_header", ""]
    dt = np.dtype([(n, codes[t]) for t, n in props])
    data = np.array([tuple(r) for r in rows], dtype=dt)
    return "\n".join(head).encode() + data.tobytes()


BASE = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
        "rot_0", "rot_1", "rot_2", "rot_3"]


def fixture(rows, sh=False, normals=False, roughness=False, metallic=False):
    """binaryPlyFixture (tests/localPly.test.ts:200-262): all-float properties."""
    names = list(BASE)
    if sh:
        names += [f"f_rest_{i}" for i in range(9)]
    if normals:
        names += ["nx", "ny", "nz"]
    if roughness:
        names.append("roughness")
    if metallic:
        names.append("metallic")
    vals = []
    for r in rows:
        v = [*r["position"], *r["dc"], r["opacity"], *r["scales"], *r["rotation"]]
        if sh:
            v += r["shRest"]
        if normals:
            v += r.get("normal", [0, 0, 0])
        if roughness:
            v.append(r.get("roughness", 0.5))
        if metallic:
            v.append(r.get("metallic", 0.0))
        vals.append(v)
    return ply_bytes([("float", n) for n in names], vals)


def f32(x):
    return np.array(x, np.float32)


def test_ply_known_answers():
    # tests/localPly.test.ts:8-89
    data = fixture([
        {"position": [1, 2, 3], "dc": [0, 1, -1], "opacity": 0, "scales": [0, math.log(2), math.log(0.5)],
         "rotation": [1, 0, 0, 0], "shRest": [0.25, 0.5, 0.75, 1, 1.25, 1.5, -0.25, -0.5, -0.75]},
        {"position": [-1, 0, 2], "dc": [2, -2, 0.5], "opacity": -2,
         "scales": [math.log(0.25), math.log(0.75), math.log(1.5)], "rotation": [0.5, 0.5, 0.5, 0.5],
         "shRest": [2, 2.25, 2.5, 3, 3.25, 3.5, 4, 4.25, 4.5]},
    ], sh=True)
    d = assets.decodeLocalPlySplatPayload("dropped.ply", data)
    assert d.count == 2 and d.sourceKind == "scaniverse_ply"
    np.testing.assert_array_equal(d.positions, f32([1, 2, 3, -1, 0, 2]))
    np.testing.assert_array_equal(d.colors, f32([0.5, 0.5 + SH_C0_TEST, 0.5 - SH_C0_TEST, 1, 0,
                                                 0.5 + SH_C0_TEST * 0.5]))
    np.testing.assert_array_equal(d.opacities, f32([0.5, 1 / (1 + math.exp(2))]))
    np.testing.assert_array_equal(d.radii, f32([2, 1.5]))
    np.testing.assert_array_equal(d.scales, f32([0, math.log(2), math.log(0.5), math.log(0.25), math.log(0.75),
                                                 math.log(1.5)]))
    np.testing.assert_array_equal(d.rotations, f32([1, 0, 0, 0, 0.5, 0.5, 0.5, 0.5]))
    assert d.originalIds.tolist() == [0, 1]
    assert d.sh["degree"] == 1 and d.sh["coefficientCount"] == 3
    np.testing.assert_array_equal(d.sh["coefficients"], f32([0.25, 1, -0.25, 0.5, 1.25, -0.5, 0.75, 1.5, -0.75,
                                                             2, 3, 4, 2.25, 3.25, 4.25, 2.5, 3.5, 4.5]))
    assert d.bounds["min"] == [-1, 0, 2] and d.bounds["max"] == [1, 2, 3] and d.bounds["center"] == [0, 1, 2.5]


def test_ply_rejects_non_ply_and_point_clouds():
    # tests/localPly.test.ts:91-124
    with pytest.raises(assets.AssetError, match="PLY header"):
        assets.decodeLocalPlySplatPayload("bad.txt", b"not ply")
    text = "\n".join(["ply", "format ascii 1.0", "comment generated_by world-tracing scripts/kaminos_wt_object.py",
                      "element vertex 2", "property float x", "property float y", "property float z",
                      "property uchar red", "property uchar green", "property uchar blue", "property uchar layer",
                      "end_header", "0.01600226 -0.59024471 0.13915212 102 103 104 0",
                      "0.00055143 -0.58335477 0.12435700 75 77 79 1", ""]).encode()
    with pytest.raises(assets.AssetError) as e:
        assets.decodeLocalPlySplatPayload("ablate02-01-denser-interior_flood-alpha_wt_layers.ply", text)
    assert str(e.value).startswith(
        "Unsupported PLY schema for hybrid splat renderer: parser=ply-point-cloud-v0; format=ascii 1.0; "
        "vertexCount=2; colorSource=rgb; properties=x,y,z,red,green,blue,layer; "
        "file=ablate02-01-denser-interior_flood-alpha_wt_layers.ply")


def test_ply_optional_material_fields():
    # tests/localPly.test.ts:126-178
    row = {"dc": [0, 0, 0], "opacity": 0, "scales": [0, 0, 0], "rotation": [1, 0, 0, 0]}
    data = fixture([dict(row, position=[1, 2, 3], normal=[0, 1, 0], roughness=0.3, metallic=0.9),
                    dict(row, position=[4, 5, 6], dc=[1, 1, 1], normal=[0.577, 0.577, 0.577], roughness=0.8,
                         metallic=0.0)], normals=True, roughness=True, metallic=True)
    d = assets.decodeLocalPlySplatPayload("baked.ply", data)
    np.testing.assert_allclose(d.normals, [0, 1, 0, 0.577, 0.577, 0.577], atol=1e-3)
    np.testing.assert_allclose(d.roughness, [0.3, 0.8], atol=1e-3)
    np.testing.assert_allclose(d.metalness, [0.9, 0.0], atol=1e-3)
    d = assets.decodeLocalPlySplatPayload("basic.ply", fixture([dict(row, position=[1, 2, 3])]))
    assert d.normals is None and d.roughness is None and d.metalness is None


def test_ply_sh3_round_trip_against_numpy():
    """A 3DGS SH3 PLY (62 float properties, f_rest channel-major) of 500 random rows: every array
    against an independent float64 numpy decode rounded to float32 once."""
    rng = np.random.default_rng(3)
    n = 500
    pos = rng.normal(0, 1, (n, 3)).astype(np.float32)
    dc = rng.normal(0, 1.5, (n, 3)).astype(np.float32)
    op = rng.normal(0, 3, n).astype(np.float32)
    sc = rng.normal(-4, 1, (n, 3)).astype(np.float32)
    rot = rng.normal(0, 1, (n, 4)).astype(np.float32)
    rest = rng.normal(0, 0.1, (n, 45)).astype(np.float32)
    props = [("float", nm) for nm in ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]]
    props += [("float", f"f_rest_{i}") for i in range(45)]
    props += [("float", nm) for nm in ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
                                       "rot_3"]]
    nrm = rng.normal(0, 1, (n, 3)).astype(np.float32)
    rows = np.concatenate([pos, nrm, dc, rest, op[:, None], sc, rot], axis=1).tolist()
    d = assets.decodeLocalPlySplatPayload("sh3.ply", ply_bytes(props, rows))
    sh_c0 = 0.28209479177387814
    np.testing.assert_array_equal(d.positions, pos.reshape(-1))
    np.testing.assert_array_equal(d.colors, np.clip(0.5 + sh_c0 * dc.astype(np.float64), 0, 1).astype(np.float32)
                                  .reshape(-1))
    np.testing.assert_array_equal(d.opacities, (1 / (1 + np.exp(-op.astype(np.float64)))).astype(np.float32))
    np.testing.assert_array_equal(d.scales, sc.reshape(-1))
    np.testing.assert_array_equal(d.radii, np.exp(sc.astype(np.float64)).max(axis=1).astype(np.float32))
    np.testing.assert_array_equal(d.rotations, rot.reshape(-1))
    np.testing.assert_array_equal(d.normals, nrm.reshape(-1))
    from hdsr.scenes import sh_from_ply_f_rest
    assert d.sh["degree"] == 3 and d.sh["coefficientCount"] == 15
    np.testing.assert_array_equal(d.sh["coefficients"], sh_from_ply_f_rest(rest).reshape(-1))
    # framing bounds: 5th / 95th percentile of the (here all 500) strided samples
    s = np.sort(pos, axis=0)
    lo, hi = int(math.floor(500 * 0.05)), min(int(math.floor(500 * (1 - 0.05))), 499)
    assert d.bounds["min"] == s[lo].astype(np.float64).tolist() and d.bounds["max"] == s[hi].astype(np.float64).tolist()
    assert d.bounds["center"] == ((s[lo].astype(np.float64) + s[hi]) * 0.5).tolist()


def test_ply_property_types_and_byte_colors():
    props = [("double", "x"), ("double", "y"), ("double", "z"), ("uchar", "red"), ("uchar", "green"),
             ("uchar", "blue"), ("short", "opacity"), ("char", "scale_0"), ("char", "scale_1"), ("char", "scale_2")]
    d = assets.decodeLocalPlySplatPayload("t.ply", ply_bytes(props, [(0.1, 1e-40, -3.5, 255, 0, 128, -1, -2, 0, 1)]))
    np.testing.assert_array_equal(d.positions, f32([0.1, 1e-40, -3.5]))
    np.testing.assert_array_equal(d.colors, f32([1.0, 0.0, 128 / 255]))
    np.testing.assert_array_equal(d.opacities, f32([1 / (1 + math.e)]))
    np.testing.assert_array_equal(d.rotations, f32([1, 0, 0, 0]))  # no rot_* -> identity
    np.testing.assert_array_equal(d.radii, f32([math.e]))
    # bounds of <= 100 rows: the exact extents of the double values
    assert d.bounds["min"] == [0.1, 1e-40, -3.5]
    # no colour properties at all: 0.5 grey; no opacity: 1
    d = assets.decodeLocalPlySplatPayload("t.ply", ply_bytes([("float", "x"), ("float", "y"), ("float", "z")],
                                                             [(1, 2, 3)]))
    np.testing.assert_array_equal(d.colors, f32([0.5, 0.5, 0.5]))
    np.testing.assert_array_equal(d.opacities, f32([1]))


@pytest.mark.parametrize("props,rows,fmt,msg", [
    ([("float", "x"), ("float", "y")], [(1, 2)], None, "PLY vertex property z is required"),
    ([("float", "x"), ("float", "y"), ("float", "z")] + [("float", f"f_rest_{i}") for i in range(4)],
     [(0,) * 7], None, "PLY f_rest_* field count must be divisible by 3 RGB channels, got 4"),
    ([("float", "x"), ("float", "y"), ("float", "z")] + [("float", f"f_rest_{i}") for i in range(6)],
     [(0,) * 9], None, "PLY f_rest_* field count 6 does not describe a complete SH degree"),
    ([("float", "x"), ("float", "y"), ("float", "z")], [(0, 0, 0)], "binary_big_endian 1.0",
     "Only binary_little_endian PLY drag-drop is supported"),
])
def test_ply_error_messages(props, rows, fmt, msg):
    data = ply_bytes(props, rows, **({"fmt": fmt} if fmt else {}))
    with pytest.raises(assets.AssetError, match="^" + msg.replace("*", r"\*") + "$"):
        assets.decodeLocalPlySplatPayload("e.ply", data)


def test_ply_header_errors():
    good = ply_bytes([("float", "x"), ("float", "y"), ("float", "z")], [(0, 0, 0), (1, 1, 1)])
    with pytest.raises(assets.AssetError, match="^t.ply ended before all 2 vertex rows could be read$"):
        assets.decodeLocalPlySplatPayload("t.ply", good[:-1])
    bad = good.replace(b"element vertex 2", b"element vertex 0")
    with pytest.raises(assets.AssetError, match="^PLY vertex count must be a positive integer$"):
        assets.decodeLocalPlySplatPayload("t.ply", bad)
    bad = good.replace(b"property float x", b"property list uchar int x")
    with pytest.raises(assets.AssetError, match="list properties are not supported"):
        assets.decodeLocalPlySplatPayload("t.ply", bad)
    bad = good.replace(b"property float x", b"property half x")
    with pytest.raises(assets.AssetError, match="^Unsupported PLY scalar property type: half$"):
        assets.decodeLocalPlySplatPayload("t.ply", bad)
    with pytest.raises(assets.AssetError, match="^Dropped file does not start with a PLY header$"):
        assets.decodeLocalPlySplatPayload("t.ply", b"PLY\n" + good[4:])
    # CRLF header lines are accepted
    cut = good.index(b"end_header\n") + len(b"end_header\n")
    d = assets.decodeLocalPlySplatPayload("t.ply", good[:cut].replace(b"\n", b"\r\n") + good[cut:])
    assert d.count == 2 and d.positions.tolist() == [0, 0, 0, 1, 1, 1]


def _raw(path, dtype, comps):
    return np.fromfile(os.path.join(os.path.dirname(SMOKE), path), dtype=dtype).reshape(-1, comps)


def test_smoke_subset_manifest_decodes_exactly():
    a = assets.loadFirstSmokeSplatPayload(SMOKE)
    man = json.load(open(SMOKE))
    rows = _raw(man["payload"]["path"], "<f4", 8)
    assert a.count == man["splat_count"] == 11801 and a.sourceKind == "scaniverse_ply"
    np.testing.assert_array_equal(a.positions, rows[:, 0:3].reshape(-1))
    np.testing.assert_array_equal(a.colors, rows[:, 3:6].reshape(-1))
    np.testing.assert_array_equal(a.opacities, rows[:, 6])
    np.testing.assert_array_equal(a.radii, rows[:, 7])
    np.testing.assert_array_equal(a.scales, _raw(man["shape"]["scales_path"], "<f4", 3).reshape(-1))
    np.testing.assert_array_equal(a.rotations, _raw(man["shape"]["rotations_path"], "<f4", 4).reshape(-1))
    np.testing.assert_array_equal(a.roughness, _raw(man["materials"]["roughness_path"], "<f4", 1).reshape(-1))
    assert a.originalIds.tolist() == list(range(11801))
    assert a.bounds["radius"] == man["bounds"]["radius"] and a.sh is None
    up = assets.uploadSplatAttributeBuffers(None, a)
    assert up["shDataBuffer"].size == 11801 * 6 and up["shDegree"] == 0  # dc + emissive, splats.ts:471-500
    assert up["materialBuffer"].dtype == np.uint32 and up["normalBuffer"] is None


def _manifest_case(tmp_path, edit=None, edit_bytes=None):
    d = tmp_path / "m"
    shutil.copytree(os.path.dirname(SMOKE), d)
    p = d / os.path.basename(SMOKE)
    man = json.load(open(p))
    if edit:
        edit(man)
    json.dump(man, open(p, "w"))
    if edit_bytes:
        edit_bytes(d, man)
    return str(p)


def _set_row_value(d, man, row, col, value):
    path = d / man["payload"]["path"]
    a = np.fromfile(path, "<f4").reshape(-1, 8)
    a[row, col] = value
    a.tofile(path)


@pytest.mark.parametrize("edit,edit_bytes,msg", [
    (lambda m: m.update(stride_bytes=36), None, "layout.strideBytes must be 32"),
    (lambda m: m.update(endianness="big"), None, "manifest.endianness must be little"),
    (lambda m: m["payload"].update(byte_length=123), None, "payload.byte_length must be 377632 for 11801 rows"),
    (lambda m: m.pop("bounds"), None, "bounds must be an object"),
    (lambda m: m["fields"][1].update(byte_offset=16), None, "layout.fields.color must be float32x3 at byte 12"),
    (None, lambda d, m: _set_row_value(d, m, 7, 6, 1.5), r"payload.rows\[7\].opacity must be in \[0, 1\]"),
    (None, lambda d, m: _set_row_value(d, m, 3, 7, 0.0), r"payload.rows\[3\].radius must be greater than zero"),
    (None, lambda d, m: _set_row_value(d, m, 5, 6, float("nan")), r"payload.rows\[5\].opacity must be a finite number"),
    (None, lambda d, m: np.array([0, 1, 1] + list(range(3, 11801)), "<u4").tofile(d / m["ids"]["path"]),
     r"originalIds\[2\] duplicates original ID 1"),
    (None, lambda d, m: np.zeros(7, "<f4").tofile(d / m["shape"]["scales_path"]),
     "shape.scales byte length 28 does not match expected 141612"),
    (lambda m: m.update(sh={"degree": 1, "basis": "3dgs_real_sh", "coefficients_layout": "splat_coeff_rgb",
                            "coefficients_component_type": "float32", "coefficient_count": 4,
                            "coefficients_byte_length": 1, "coefficients_path": "x.bin"}), None,
     "sh.coefficient_count must be 3 for degree 1"),
])
def test_manifest_validation_messages(tmp_path, edit, edit_bytes, msg):
    p = _manifest_case(tmp_path, edit, edit_bytes)
    if "sh" in json.load(open(p)):
        open(os.path.join(os.path.dirname(p), "x.bin"), "wb").write(b"\0" * 4)
    with pytest.raises(assets.AssetError, match="^" + msg + "$"):
        assets.loadFirstSmokeSplatPayload(p)


def test_inline_payload_rows_and_planar():
    bounds = {"min": [0, 0, 0], "max": [1, 1, 1], "center": [0.5, 0.5, 0.5], "radius": 0.9}
    lay = {"strideBytes": 32, "fields": assets.FIRST_SMOKE_SPLAT_LAYOUT["fields"]}
    rows = {"count": 2, "bounds": bounds, "layout": lay,
            "splats": [{"position": [0, 1, 2], "color": [1, 0, 0], "opacity": 0.5, "radius": 0.25},
                       {"position": [3, 4, 5], "color": [0, 1, 0], "opacity": 1, "radius": 2}]}
    a = assets.decodeFirstSmokeSplatPayload(rows)
    np.testing.assert_array_equal(a.scales, f32([math.log(0.25)] * 3 + [math.log(2)] * 3))
    np.testing.assert_array_equal(a.rotations, f32([1, 0, 0, 0, 1, 0, 0, 0]))
    planar = {"metadata": {"count": 2, "bounds": bounds, "layout": lay},
              "attributes": {"positions": [0, 1, 2, 3, 4, 5], "colors": [1, 0, 0, 0, 1, 0], "opacities": [0.5, 1],
                             "radii": [0.25, 2]}, "originalIds": [1, 0]}
    b = assets.decodeFirstSmokeSplatPayload(planar)
    np.testing.assert_array_equal(b.positions, a.positions)
    assert b.originalIds.tolist() == [1, 0]
    bad = dict(planar, originalIds=[1, 2])
    with pytest.raises(assets.AssetError, match=r"^originalIds\[1\] must be less than count 2$"):
        assets.decodeFirstSmokeSplatPayload(bad)
