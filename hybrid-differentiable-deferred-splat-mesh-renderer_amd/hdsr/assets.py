"""On-disk splat assets -> SplatAttributes -> the compositor's input buffers.

Mirrors the reference's loaders, same names, argument meaning and error messages:
  decodeLocalPlySplatPayload (src/localPly.ts:272-420)      binary little-endian 3DGS / Scaniverse
                                                            PLY, decoded by the library's C++
                                                            decoder (hdsr_ply_decode)
  decodeFirstSmokeSplatManifest (src/splats.ts:230-344)     JSON manifest + little-endian float32 /
                                                            uint32 sidecars
  decodeFirstSmokeSplatPayload (src/splats.ts:203-228)      inline JSON payload (rows or planar)
  loadFirstSmokeSplatPayload                                fetchFirstSmokeSplatPayload
                                                            (src/splats.ts:114-201) over files: a
                                                            sidecar manifest or an inline payload
  uploadSplatAttributeBuffers (src/splats.ts:376-428)       the packers: buffers for
                                                            createTileSplatBindGroups
SplatAttributes keeps the reference's field names (src/splats.ts:38-60); arrays are flat numpy
arrays as the reference's typed arrays are.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .scenes import Scene, pack_material, pack_sh_data

FIRST_SMOKE_SPLAT_LAYOUT = {
    "strideBytes": 32,
    "fields": [
        {"name": "position", "type": "float32", "components": 3, "byteOffset": 0},
        {"name": "color", "type": "float32", "components": 3, "byteOffset": 12},
        {"name": "opacity", "type": "float32", "components": 1, "byteOffset": 24},
        {"name": "radius", "type": "float32", "components": 1, "byteOffset": 28},
    ],
}


@dataclass
class SplatAttributes:
    count: int
    sourceKind: str
    positions: np.ndarray
    colors: np.ndarray
    opacities: np.ndarray
    radii: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray
    originalIds: np.ndarray
    bounds: dict
    layout: dict
    sh: dict | None = None
    normals: np.ndarray | None = None
    roughness: np.ndarray | None = None
    metalness: np.ndarray | None = None
    emissive: np.ndarray | None = None
    detailNormals: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


class AssetError(ValueError):
    """A decode failure, with the reference's message."""


# ----------------------------------------------------------------------------------------------
# PLY (C++ decoder in the library)

_PlyInfo = _lib.PlyInfo
_PlyOut = _lib.SplatAttributesOut


PLY_NORMALS, PLY_ROUGHNESS, PLY_METALNESS, PLY_EMISSIVE, PLY_DETAIL_NORMALS = 1, 2, 4, 8, 16


def _ply_lib():
    return _lib.load()


def decodeLocalPlySplatPayload(fileName: str, data) -> SplatAttributes:
    """decodeLocalPlySplatPayload (src/localPly.ts:272-420) over the library's decoder."""
    buf = np.frombuffer(memoryview(data), dtype=np.uint8)
    lib = _ply_lib()
    err = C.create_string_buffer(1024)
    name = fileName.encode()
    info = _PlyInfo()
    rc = lib.hdsr_ply_inspect(name, buf.ctypes.data, buf.size, C.byref(info), err, len(err))
    if rc:
        raise AssetError(err.value.decode())
    n, k, f = info.count, info.sh_coefficient_count, info.fields
    z = lambda *s: np.zeros(s, np.float32)  # noqa: E731
    arrays = {"positions": z(n * 3), "colors": z(n * 3), "opacities": z(n), "radii": z(n), "scales": z(n * 3),
              "rotations": z(n * 4), "original_ids": np.zeros(n, np.uint32),
              "sh_coefficients": z(n * k * 3) if k else None,
              "normals": z(n * 3) if f & PLY_NORMALS else None, "roughness": z(n) if f & PLY_ROUGHNESS else None,
              "metalness": z(n) if f & PLY_METALNESS else None, "emissive": z(n * 3) if f & PLY_EMISSIVE else None,
              "detail_normals": z(n * 3) if f & PLY_DETAIL_NORMALS else None}
    out = _PlyOut(*[None if arrays[nm] is None else arrays[nm].ctypes.data for nm, _ in _PlyOut._fields_[:13]])
    rc = lib.hdsr_ply_decode(name, buf.ctypes.data, buf.size, C.byref(out), err, len(err))
    if rc:
        raise AssetError(err.value.decode())
    bounds = {"min": list(out.bounds_min), "max": list(out.bounds_max), "center": list(out.bounds_center),
              "radius": float(out.bounds_radius)}
    sh = None
    if k:
        sh = {"degree": info.sh_degree, "basis": "3dgs_real_sh", "coefficientCount": k, "layout": "splat_coeff_rgb",
              "coefficients": arrays["sh_coefficients"]}
    return SplatAttributes(n, "scaniverse_ply", arrays["positions"], arrays["colors"], arrays["opacities"],
                           arrays["radii"], arrays["scales"], arrays["rotations"], arrays["original_ids"], bounds,
                           _copy_layout(FIRST_SMOKE_SPLAT_LAYOUT), sh, arrays["normals"], arrays["roughness"],
                           arrays["metalness"], arrays["emissive"], arrays["detail_normals"])


# ----------------------------------------------------------------------------------------------
# first-smoke manifest / payload (src/splats.ts)

_UNDEF = object()


def _get(d: dict, k):
    return d[k] if k in d else _UNDEF


def _first_defined(a, b):
    return b if a is _UNDEF else a


def _is_number(v) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _require_record(v, path: str) -> dict:
    if not isinstance(v, dict):
        raise AssetError(f"{path} must be an object")
    return v


def _require_finite(v, path: str) -> float:
    if not _is_number(v) or not math.isfinite(v):
        raise AssetError(f"{path} must be a finite number")
    return v


def _require_string(v, path: str) -> str:
    if not isinstance(v, str) or len(v) == 0:
        raise AssetError(f"{path} must be a non-empty string")
    return v


def _require_positive_finite(v, path: str) -> float:
    x = _require_finite(v, path)
    if x <= 0:
        raise AssetError(f"{path} must be greater than zero")
    return x


def _require_unit(v, path: str) -> float:
    x = _require_finite(v, path)
    if x < 0 or x > 1:
        raise AssetError(f"{path} must be in [0, 1]")
    return x


def _require_nonneg_int(v, path: str) -> int:
    if not _is_number(v) or not math.isfinite(v) or float(v) != math.floor(v) or v < 0:
        raise AssetError(f"{path} must be a non-negative integer")
    return int(v)


def _require_positive_int(v, path: str) -> int:
    x = _require_nonneg_int(v, path)
    if x <= 0:
        raise AssetError(f"{path} must be greater than zero")
    return x


def _require_vec3(v, path: str) -> list:
    if not isinstance(v, list) or len(v) != 3:
        raise AssetError(f"{path} must be a float3")
    return [_require_finite(v[i], f"{path}[{i}]") for i in range(3)]


def _copy_layout(layout: dict) -> dict:
    return {"strideBytes": layout["strideBytes"], "fields": [dict(f) for f in layout["fields"]]}


def _decode_bounds(v, path: str) -> dict:
    b = _require_record(None if v is _UNDEF else v, path)
    return {"min": _require_vec3(b.get("min"), f"{path}.min"), "max": _require_vec3(b.get("max"), f"{path}.max"),
            "center": _require_vec3(b.get("center"), f"{path}.center"),
            "radius": _require_positive_finite(b.get("radius"), f"{path}.radius")}


def _decode_field_layout(v, path: str) -> dict:
    f = _require_record(v, path)
    name = f.get("name")
    if name not in ("position", "color", "opacity", "radius"):
        raise AssetError(f"{path}.name is not a first-smoke row field")
    if _first_defined(_get(f, "type"), _get(f, "component_type")) != "float32":
        raise AssetError(f"{path}.type must be float32")
    comps = _require_positive_int(f.get("components"), f"{path}.components")
    if comps not in (1, 3):
        raise AssetError(f"{path}.components must be 1 or 3")
    off = _require_nonneg_int(_first_defined(_get(f, "byteOffset"), _get(f, "byte_offset")), f"{path}.byteOffset")
    return {"name": name, "type": "float32", "components": comps, "byteOffset": off}


def _layout_source(meta: dict):
    if "layout" in meta:
        return meta["layout"]
    if any(k in meta for k in ("strideBytes", "stride_bytes", "fields", "fieldLayout")):
        stride = _first_defined(_get(meta, "strideBytes"), _get(meta, "stride_bytes"))
        fields = _first_defined(_get(meta, "fields"), _get(meta, "fieldLayout"))
        return {"strideBytes": None if stride is _UNDEF else stride, "fields": None if fields is _UNDEF else fields}
    return _UNDEF


def _decode_layout(v, path: str) -> dict:
    layout = _require_record(None if v is _UNDEF else v, path)
    stride = _require_positive_int(layout.get("strideBytes"), f"{path}.strideBytes")
    if stride != FIRST_SMOKE_SPLAT_LAYOUT["strideBytes"]:
        raise AssetError(f"{path}.strideBytes must be {FIRST_SMOKE_SPLAT_LAYOUT['strideBytes']}")
    fields = layout.get("fields")
    if not isinstance(fields, list):
        raise AssetError(f"{path}.fields must be an array")
    if len(fields) != len(FIRST_SMOKE_SPLAT_LAYOUT["fields"]):
        raise AssetError(f"{path}.fields must contain exactly the first-smoke row fields")
    decoded = [_decode_field_layout(f, f"{path}.fields[{i}]") for i, f in enumerate(fields)]
    for req in FIRST_SMOKE_SPLAT_LAYOUT["fields"]:
        got = next((f for f in decoded if f["name"] == req["name"]), None)
        if got is None:
            raise AssetError(f"{path}.fields is missing {req['name']}")
        if got["components"] != req["components"] or got["byteOffset"] != req["byteOffset"]:
            raise AssetError(f"{path}.fields.{req['name']} must be {req['type']}x{req['components']} at byte "
                             f"{req['byteOffset']}")
    return _copy_layout(FIRST_SMOKE_SPLAT_LAYOUT)


def _decode_float32_bytes(data, expected_len: int, path: str) -> np.ndarray:
    expected = expected_len * 4
    if len(data) != expected:
        raise AssetError(f"{path} byte length {len(data)} does not match expected {expected}")
    v = np.frombuffer(data, dtype="<f4").astype(np.float32)
    bad = np.flatnonzero(~np.isfinite(v))
    if bad.size:
        raise AssetError(f"{path}[{int(bad[0])}] must be a finite number")
    return v


def _decode_original_ids(values, count: int) -> np.ndarray:
    """decodeOriginalIds (src/splats.ts:806-834): rows in order; the first failing row raises."""
    if values is None:
        return np.arange(count, dtype=np.uint32)
    if isinstance(values, np.ndarray):  # a uint32 sidecar: integers by construction
        ids = values.astype(np.int64)
        if ids.size != count:
            raise AssetError(f"originalIds must contain {count} entries")
        over = np.flatnonzero(ids >= count)
        order = np.argsort(ids, kind="stable")
        s = ids[order]
        dups = np.flatnonzero(s[1:] == s[:-1])
        first_over = int(over[0]) if over.size else count
        first_dup = int(np.min(order[dups + 1])) if dups.size else count  # a later occurrence of a seen ID
        if first_over < count and first_over <= first_dup:
            raise AssetError(f"originalIds[{first_over}] must be less than count {count}")
        if first_dup < count:
            raise AssetError(f"originalIds[{first_dup}] duplicates original ID {int(ids[first_dup])}")
        return ids.astype(np.uint32)
    if not isinstance(values, list):
        raise AssetError("originalIds must be an array or typed array")
    if len(values) != count:
        raise AssetError(f"originalIds must contain {count} entries")
    seen = set()
    out = np.zeros(count, np.uint32)
    for i, v in enumerate(values):
        x = _require_nonneg_int(v, f"originalIds[{i}]")
        if x >= count:
            raise AssetError(f"originalIds[{i}] must be less than count {count}")
        if x in seen:
            raise AssetError(f"originalIds[{i}] duplicates original ID {x}")
        seen.add(x)
        out[i] = x
    return out


def _decode_sh_sidecar(root: dict, count: int, sh_bytes):
    if "sh" not in root:
        return None
    sh = _require_record(root["sh"], "sh")
    degree = _require_nonneg_int(sh.get("degree"), "sh.degree")
    if degree <= 0:
        raise AssetError("sh.degree must be greater than zero when an SH sidecar is declared")
    if sh.get("basis") != "3dgs_real_sh":
        raise AssetError("sh.basis must be 3dgs_real_sh")
    if sh.get("coefficients_layout") != "splat_coeff_rgb":
        raise AssetError("sh.coefficients_layout must be splat_coeff_rgb")
    if sh.get("coefficients_component_type") != "float32":
        raise AssetError("sh.coefficients_component_type must be float32")
    k = (degree + 1) ** 2 - 1
    declared = _require_positive_int(sh.get("coefficient_count"), "sh.coefficient_count")
    if declared != k:
        raise AssetError(f"sh.coefficient_count must be {k} for degree {degree}")
    expected = count * k * 3 * 4
    declared_bytes = _require_positive_int(sh.get("coefficients_byte_length"), "sh.coefficients_byte_length")
    if declared_bytes != expected:
        raise AssetError(f"sh.coefficients_byte_length must be {expected} for {count} splats at degree {degree}")
    if sh_bytes is None:
        raise AssetError("SH sidecar bytes are required when manifest.sh is declared")
    return {"degree": degree, "basis": "3dgs_real_sh", "coefficientCount": k, "layout": "splat_coeff_rgb",
            "coefficients": _decode_float32_bytes(sh_bytes, count * k * 3, "sh.coefficients")}


def _manifest_source_kind(root: dict) -> str:
    if "source" in root:
        src = _require_record(root["source"], "source")
        if "kind" in src:
            kind = _require_string(src["kind"], "source.kind")
            return "real_splat_spz" if kind == "spz" else kind
    return "real_scaniverse_ply"


def _fallback_scales(radii: np.ndarray) -> np.ndarray:
    lr = np.log(np.maximum(radii.astype(np.float64), 1e-12)).astype(np.float32)
    return np.repeat(lr, 3)


def _identity_rotations(count: int) -> np.ndarray:
    r = np.zeros(count * 4, np.float32)
    r[0::4] = 1.0
    return r


def decodeFirstSmokeSplatManifest(manifest, payloadBytes, idsBytes=None, scalesBytes=None, rotationsBytes=None,
                                  shBytes=None, roughnessBytes=None, metalnessBytes=None) -> SplatAttributes:
    """decodeFirstSmokeSplatManifest (src/splats.ts:230-344)."""
    root = _require_record(manifest, "manifest")
    count = _require_positive_int(_first_defined(_get(root, "splat_count"), _get(root, "count")), "splat_count")
    if _first_defined(_get(root, "endianness"), "little") != "little":
        raise AssetError("manifest.endianness must be little")
    bounds = _decode_bounds(_get(root, "bounds"), "bounds")
    source_kind = _manifest_source_kind(root)
    layout = _decode_layout(_layout_source(root), "layout")
    info = _require_record(root.get("payload"), "payload")
    expected = count * layout["strideBytes"]
    declared = _require_positive_int(_first_defined(_get(info, "byte_length"), expected), "payload.byte_length")
    if declared != expected:
        raise AssetError(f"payload.byte_length must be {expected} for {count} rows")
    if len(payloadBytes) != expected:
        raise AssetError(f"payload sidecar byte length {len(payloadBytes)} does not match expected {expected}")
    rows = np.frombuffer(payloadBytes, dtype="<f4").reshape(count, 8).astype(np.float32)
    positions = np.ascontiguousarray(rows[:, 0:3]).reshape(-1)
    colors = np.ascontiguousarray(rows[:, 3:6]).reshape(-1)
    op, rad = rows[:, 6].copy(), rows[:, 7].copy()
    bad_op = ~np.isfinite(op) | (op < 0) | (op > 1)
    bad_rad = ~np.isfinite(rad) | (rad <= 0)
    bad = np.flatnonzero(bad_op | bad_rad)
    if bad.size:
        r = int(bad[0])
        if bad_op[r]:
            if not np.isfinite(op[r]):
                raise AssetError(f"payload.rows[{r}].opacity must be a finite number")
            raise AssetError(f"payload.rows[{r}].opacity must be in [0, 1]")
        if not np.isfinite(rad[r]):
            raise AssetError(f"payload.rows[{r}].radius must be a finite number")
        raise AssetError(f"payload.rows[{r}].radius must be greater than zero")
    scales = _fallback_scales(rad) if scalesBytes is None else _decode_float32_bytes(scalesBytes, count * 3,
                                                                                     "shape.scales")
    rotations = _identity_rotations(count) if rotationsBytes is None else _decode_float32_bytes(
        rotationsBytes, count * 4, "shape.rotations")
    sh = _decode_sh_sidecar(root, count, shBytes)
    roughness = None if roughnessBytes is None else _decode_float32_bytes(roughnessBytes, count, "materials.roughness")
    metalness = None if metalnessBytes is None else _decode_float32_bytes(metalnessBytes, count, "materials.metalness")
    if idsBytes is None:
        ids = _decode_original_ids(None, count)
    else:
        if len(idsBytes) != count * 4:
            raise AssetError(f"ID sidecar byte length {len(idsBytes)} does not match expected {count * 4}")
        ids = _decode_original_ids(np.frombuffer(idsBytes, dtype="<u4"), count)
    return SplatAttributes(count, source_kind, positions, colors, op, rad, scales, rotations, ids, bounds, layout, sh,
                           roughness=roughness, metalness=metalness)


def _decode_float_list(v, expected_len: int, path: str, validator=_require_finite) -> np.ndarray:
    if isinstance(v, np.ndarray):
        v = v.tolist()
    if not isinstance(v, list):
        raise AssetError(f"{path} must be an array or typed array")
    if len(v) != expected_len:
        raise AssetError(f"{path} must contain {expected_len} entries")
    return np.array([validator(x, f"{path}[{i}]") for i, x in enumerate(v)], np.float32)


def decodeFirstSmokeSplatPayload(payload) -> SplatAttributes:
    """decodeFirstSmokeSplatPayload (src/splats.ts:203-228): inline rows or planar attributes."""
    root = _require_record(payload, "payload")
    meta = root if "metadata" not in root else _require_record(root["metadata"], "metadata")
    count = _require_positive_int(_first_defined(_get(meta, "count"), _get(meta, "splat_count")), "count")
    bounds = _decode_bounds(_get(meta, "bounds"), "bounds")
    layout = _decode_layout(_layout_source(meta), "layout")
    if "splats" in root:
        rows = root["splats"]
        if not isinstance(rows, list):
            raise AssetError("splats must be an array of first-smoke splat rows")
        if len(rows) != count:
            raise AssetError(f"count {count} does not match {len(rows)} rows")
        pos, col, op, rad = [], [], [], []
        for i, r in enumerate(rows):
            row = _require_record(r, f"splats[{i}]")
            pos += _require_vec3(row.get("position"), f"splats[{i}].position")
            col += _require_vec3(row.get("color"), f"splats[{i}].color")
            op.append(_require_unit(row.get("opacity"), f"splats[{i}].opacity"))
            rad.append(_require_positive_finite(row.get("radius"), f"splats[{i}].radius"))
        positions, colors = np.array(pos, np.float32), np.array(col, np.float32)
        opacities, radii = np.array(op, np.float32), np.array(rad, np.float32)
        scales, rotations = _fallback_scales(radii), _identity_rotations(count)
    else:
        a = _require_record(root.get("attributes"), "attributes")
        positions = _decode_float_list(_first_defined(_get(a, "positions"), _get(a, "position")), count * 3,
                                       "attributes.positions")
        colors = _decode_float_list(_first_defined(_get(a, "colors"), _get(a, "color")), count * 3,
                                    "attributes.colors")
        opacities = _decode_float_list(_first_defined(_get(a, "opacities"), _get(a, "opacity")), count,
                                       "attributes.opacities", _require_unit)
        radii = _decode_float_list(_first_defined(_get(a, "radii"), _get(a, "radius")), count, "attributes.radii",
                                   _require_positive_finite)
        scales = _fallback_scales(radii) if "scales" not in a else _decode_float_list(a["scales"], count * 3,
                                                                                        "attributes.scales")
        rotations = _identity_rotations(count) if "rotations" not in a else _decode_float_list(
            a["rotations"], count * 4, "attributes.rotations")
    src = _first_defined(_get(meta, "sourceKind"), "real_scaniverse_ply")
    ids = _decode_original_ids(root.get("originalIds"), count)
    return SplatAttributes(count, _require_string(src, "sourceKind"), positions, colors, opacities, radii, scales,
                           rotations, ids, bounds, layout)


def _resolve_sidecar(manifest_path: str, path: str) -> str:
    if os.path.isabs(path):
        return path
    return os.path.join(os.path.dirname(manifest_path), path)


def _read(path: str) -> bytes:
    with open(path, "rb") as f:
        return f.read()


def loadFirstSmokeSplatPayload(path: str) -> SplatAttributes:
    """fetchFirstSmokeSplatPayload (src/splats.ts:114-201) over local files: a sidecar manifest
    (its payload / ids / shape / SH / material sidecars resolved next to it) or an inline payload."""
    with open(path, "r", encoding="utf-8") as f:
        payload = json.load(f)
    is_manifest = (isinstance(payload, dict) and isinstance(payload.get("payload"), dict)
                   and isinstance(payload["payload"].get("path"), str))
    if not is_manifest:
        return decodeFirstSmokeSplatPayload(payload)
    root = payload
    info = _require_record(root["payload"], "payload.payload")
    side = lambda p: _read(_resolve_sidecar(path, p))  # noqa: E731
    payload_bytes = side(_require_string(info.get("path"), "payload.path"))
    ids_path = None
    if "ids" in root and "path" in _require_record(root["ids"], "ids"):
        ids_path = _require_string(root["ids"]["path"], "ids.path")
    elif "identity" in root and "ids_path" in _require_record(root["identity"], "identity"):
        ids_path = _require_string(root["identity"]["ids_path"], "identity.ids_path")
    shape = _require_record(root["shape"], "shape") if "shape" in root else {}
    scales_path = _require_string(shape["scales_path"], "shape.scales_path") if "scales_path" in shape else None
    rot_path = _require_string(shape["rotations_path"], "shape.rotations_path") if "rotations_path" in shape else None
    sh_path = None
    if "sh" in root:
        sh_path = _require_string(_require_record(root["sh"], "sh").get("coefficients_path"), "sh.coefficients_path")
    mats = root.get("materials") or {}
    rough_path, metal_path = mats.get("roughness_path"), mats.get("metalness_path")
    opt = lambda p: None if p is None else side(p)  # noqa: E731
    return decodeFirstSmokeSplatManifest(payload, payload_bytes, opt(ids_path), opt(scales_path), opt(rot_path),
                                         opt(sh_path), opt(rough_path), opt(metal_path))


def framingFromBounds(bounds: dict, options: dict | None = None) -> dict:
    """framingFromBounds (src/splats.ts:345-371): camera target, distance and clip planes."""
    o = options or {}
    fov_y = o.get("fovY", math.pi / 3)
    padding = o.get("padding", 1.35)
    min_near = o.get("minNear", 0.001)
    if not math.isfinite(fov_y) or fov_y <= 0 or fov_y >= math.pi:
        raise AssetError("framing.fovY must be a finite angle in (0, pi)")
    if not math.isfinite(padding) or padding < 1:
        raise AssetError("framing.padding must be a finite value >= 1")
    d = [bounds["max"][i] - bounds["min"][i] for i in range(3)]
    radius = max(bounds["radius"], math.hypot(*d) / 2)
    padded = radius * padding
    distance = padded / math.sin(fov_y / 2)
    return {"target": list(bounds["center"]), "radius": radius, "distance": distance,
            "near": max(min_near, distance - padded * 2), "far": distance + padded * 2}


# ----------------------------------------------------------------------------------------------
# packers (src/splats.ts:376-500)

def packNormalBuffer(count: int, normals: np.ndarray, detailNormals: np.ndarray | None = None) -> np.ndarray:
    """[geometric normals (3N) | detail normals (3N)] (src/splats.ts:448-454)."""
    if detailNormals is None:
        return np.asarray(normals, np.float32)
    return np.concatenate([np.asarray(normals, np.float32).reshape(-1), np.asarray(detailNormals, np.float32)])


def uploadSplatAttributeBuffers(device, attributes: SplatAttributes) -> dict:
    """uploadSplatAttributeBuffers (src/splats.ts:376-428): the host buffers that
    createTileSplatBindGroups binds (the library copies or borrows them at bind time)."""
    a = attributes
    n = a.count
    k = a.sh["coefficientCount"] if a.sh else 0
    sh = None if not a.sh else np.asarray(a.sh["coefficients"], np.float32).reshape(n, k, 3)
    em = None if a.emissive is None else np.asarray(a.emissive, np.float32).reshape(n, 3)
    return {
        "count": n,
        "positionBuffer": np.asarray(a.positions, np.float32),
        "opacityBuffer": np.asarray(a.opacities, np.float32),
        "scaleBuffer": np.asarray(a.scales, np.float32),
        "rotationBuffer": np.asarray(a.rotations, np.float32),
        "materialBuffer": pack_material(n, a.roughness, a.metalness),
        "normalBuffer": None if a.normals is None else packNormalBuffer(n, a.normals, a.detailNormals),
        "shDataBuffer": pack_sh_data(np.asarray(a.colors, np.float32).reshape(n, 3), sh, em),
        "shDegree": a.sh["degree"] if a.sh else 0,
        "originalIdBuffer": np.asarray(a.originalIds, np.uint32),
        "bounds": a.bounds,
        "layout": a.layout,
    }


def scene_from_attributes(attributes: SplatAttributes) -> Scene:
    """A Scene (the tests' and benchmark's bind form) holding exactly the uploaded buffers."""
    a = attributes
    n = a.count
    k = a.sh["coefficientCount"] if a.sh else 0
    sh = None if not a.sh else np.asarray(a.sh["coefficients"], np.float32).reshape(n, k, 3)
    normals = None if a.normals is None else packNormalBuffer(n, a.normals, a.detailNormals)
    return Scene(np.asarray(a.positions, np.float32).reshape(n, 3), np.asarray(a.scales, np.float32).reshape(n, 3),
                 np.asarray(a.rotations, np.float32).reshape(n, 4), np.asarray(a.opacities, np.float32),
                 np.asarray(a.colors, np.float32).reshape(n, 3), sh, a.sh["degree"] if a.sh else 0,
                 pack_material(n, a.roughness, a.metalness), normals,
                 None if a.emissive is None else np.asarray(a.emissive, np.float32).reshape(n, 3),
                 {"sourceKind": a.sourceKind, "bounds": a.bounds})
