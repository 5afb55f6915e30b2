'use strict';
// splatAssets.js - the reference's splat asset loaders for Node (>= 12), same names and
// messages, feeding createTileSplatBindGroups:
//   decodeLocalPlySplatPayload(fileName, bytes)        src/localPly.ts:272-420, over the library's
//                                                      C++ decoder (hdsr_ply_decode)
//   decodeFirstSmokeSplatManifest(manifest, payload,
//     ids?, scales?, rotations?, sh?, rough?, metal?)  src/splats.ts:230-344
//   decodeFirstSmokeSplatPayload(payload)              src/splats.ts:203-228
//   fetchFirstSmokeSplatPayload(path)                  src/splats.ts:114-201 over the file system
//                                                      (a Promise; loadFirstSmokeSplatPayload is
//                                                      the synchronous form)
//   uploadSplatAttributeBuffers(device, attributes)    src/splats.ts:376-500: the typed arrays
//                                                      createTileSplatBindGroups binds
//   framingFromBounds(bounds, options?)                src/splats.ts:345-371
var fs = require('fs');
var path = require('path');
var addon = require('./hdsr_napi.node');

var FIRST_SMOKE_SPLAT_LAYOUT = {
  strideBytes: 32,
  fields: [
    { name: 'position', type: 'float32', components: 3, byteOffset: 0 },
    { name: 'color', type: 'float32', components: 3, byteOffset: 12 },
    { name: 'opacity', type: 'float32', components: 1, byteOffset: 24 },
    { name: 'radius', type: 'float32', components: 1, byteOffset: 28 },
  ],
};

function copyLayout(l) {
  return { strideBytes: l.strideBytes, fields: l.fields.map(function (f) { return Object.assign({}, f); }) };
}

function decodeLocalPlySplatPayload(fileName, bytes) {
  var d = addon.plyDecode(String(fileName), bytes);
  return {
    count: d.count,
    sourceKind: 'scaniverse_ply',
    positions: d.positions,
    colors: d.colors,
    opacities: d.opacities,
    radii: d.radii,
    scales: d.scales,
    rotations: d.rotations,
    sh: d.shCoefficients === undefined ? undefined : {
      degree: d.shDegree, basis: '3dgs_real_sh', coefficientCount: d.shCoefficientCount,
      layout: 'splat_coeff_rgb', coefficients: d.shCoefficients,
    },
    normals: d.normals,
    roughness: d.roughness,
    metalness: d.metalness,
    emissive: d.emissive,
    detailNormals: d.detailNormals,
    originalIds: d.originalIds,
    bounds: d.bounds,
    layout: copyLayout(FIRST_SMOKE_SPLAT_LAYOUT),
  };
}

// ---- validation helpers (src/splats.ts:836-915) ----
function requireRecord(v, p) {
  if (!v || typeof v !== 'object' || Array.isArray(v)) throw new Error(p + ' must be an object');
  return v;
}
function requireFiniteNumber(v, p) {
  if (typeof v !== 'number' || !isFinite(v)) throw new Error(p + ' must be a finite number');
  return v;
}
function requireVec3(v, p) {
  if (!Array.isArray(v) || v.length !== 3) throw new Error(p + ' must be a float3');
  return [requireFiniteNumber(v[0], p + '[0]'), requireFiniteNumber(v[1], p + '[1]'), requireFiniteNumber(v[2], p + '[2]')];
}
function requireString(v, p) {
  if (typeof v !== 'string' || v.length === 0) throw new Error(p + ' must be a non-empty string');
  return v;
}
function requirePositiveFinite(v, p) {
  var x = requireFiniteNumber(v, p);
  if (x <= 0) throw new Error(p + ' must be greater than zero');
  return x;
}
function requireUnitInterval(v, p) {
  var x = requireFiniteNumber(v, p);
  if (x < 0 || x > 1) throw new Error(p + ' must be in [0, 1]');
  return x;
}
function requireNonNegativeInteger(v, p) {
  if (typeof v !== 'number' || !Number.isInteger(v) || v < 0) throw new Error(p + ' must be a non-negative integer');
  return v;
}
function requirePositiveInteger(v, p) {
  var x = requireNonNegativeInteger(v, p);
  if (x <= 0) throw new Error(p + ' must be greater than zero');
  return x;
}
function firstDefined(a, b) { return a === undefined ? b : a; }

function decodeBounds(v, p) {
  var b = requireRecord(v, p);
  return { min: requireVec3(b.min, p + '.min'), max: requireVec3(b.max, p + '.max'),
    center: requireVec3(b.center, p + '.center'), radius: requirePositiveFinite(b.radius, p + '.radius') };
}

function decodeFieldLayout(v, p) {
  var f = requireRecord(v, p);
  var name = f.name;
  if (name !== 'position' && name !== 'color' && name !== 'opacity' && name !== 'radius') {
    throw new Error(p + '.name is not a first-smoke row field');
  }
  if (firstDefined(f.type, f.component_type) !== 'float32') throw new Error(p + '.type must be float32');
  var components = requirePositiveInteger(f.components, p + '.components');
  if (components !== 1 && components !== 3) throw new Error(p + '.components must be 1 or 3');
  var byteOffset = requireNonNegativeInteger(firstDefined(f.byteOffset, f.byte_offset), p + '.byteOffset');
  return { name: name, type: 'float32', components: components, byteOffset: byteOffset };
}

function decodeLayout(v, p) {
  var l = requireRecord(v, p);
  var stride = requirePositiveInteger(l.strideBytes, p + '.strideBytes');
  if (stride !== FIRST_SMOKE_SPLAT_LAYOUT.strideBytes) {
    throw new Error(p + '.strideBytes must be ' + FIRST_SMOKE_SPLAT_LAYOUT.strideBytes);
  }
  if (!Array.isArray(l.fields)) throw new Error(p + '.fields must be an array');
  if (l.fields.length !== FIRST_SMOKE_SPLAT_LAYOUT.fields.length) {
    throw new Error(p + '.fields must contain exactly the first-smoke row fields');
  }
  var decoded = l.fields.map(function (f, i) { return decodeFieldLayout(f, p + '.fields[' + i + ']'); });
  FIRST_SMOKE_SPLAT_LAYOUT.fields.forEach(function (req) {
    var got = decoded.filter(function (c) { return c.name === req.name; })[0];
    if (!got) throw new Error(p + '.fields is missing ' + req.name);
    if (got.type !== req.type || got.components !== req.components || got.byteOffset !== req.byteOffset) {
      throw new Error(p + '.fields.' + req.name + ' must be ' + req.type + 'x' + req.components + ' at byte ' + req.byteOffset);
    }
  });
  return copyLayout(FIRST_SMOKE_SPLAT_LAYOUT);
}

function layoutSource(meta) {
  if (meta.layout !== undefined) return meta.layout;
  if (meta.strideBytes !== undefined || meta.stride_bytes !== undefined || meta.fields !== undefined ||
      meta.fieldLayout !== undefined) {
    return { strideBytes: firstDefined(meta.strideBytes, meta.stride_bytes), fields: firstDefined(meta.fields, meta.fieldLayout) };
  }
  return undefined;
}

function decodeOriginalIds(value, count) {
  var ids = new Uint32Array(count);
  if (value === undefined) {
    for (var i = 0; i < count; i++) ids[i] = i;
    return ids;
  }
  if (!Array.isArray(value) && !ArrayBuffer.isView(value)) throw new Error('originalIds must be an array or typed array');
  var values = Array.prototype.slice.call(value);
  if (values.length !== count) throw new Error('originalIds must contain ' + count + ' entries');
  var seen = new Set();
  for (var j = 0; j < values.length; j++) {
    var id = requireNonNegativeInteger(values[j], 'originalIds[' + j + ']');
    if (id >= count) throw new Error('originalIds[' + j + '] must be less than count ' + count);
    if (seen.has(id)) throw new Error('originalIds[' + j + '] duplicates original ID ' + id);
    seen.add(id);
    ids[j] = id;
  }
  return ids;
}

function asArrayBuffer(b) {
  if (b instanceof ArrayBuffer) return b;
  return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength);
}

function decodeFloat32Bytes(bytes, expectedLength, p) {
  var ab = asArrayBuffer(bytes);
  var expected = expectedLength * 4;
  if (ab.byteLength !== expected) throw new Error(p + ' byte length ' + ab.byteLength + ' does not match expected ' + expected);
  var data = new DataView(ab);
  var out = new Float32Array(expectedLength);
  for (var i = 0; i < expectedLength; i++) out[i] = requireFiniteNumber(data.getFloat32(i * 4, true), p + '[' + i + ']');
  return out;
}

function fallbackScalesFromRadii(radii) {
  var s = new Float32Array(radii.length * 3);
  for (var i = 0; i < radii.length; i++) {
    var l = Math.log(Math.max(radii[i], 1e-12));
    s[i * 3] = l; s[i * 3 + 1] = l; s[i * 3 + 2] = l;
  }
  return s;
}

function identityRotations(count) {
  var r = new Float32Array(count * 4);
  for (var i = 0; i < count; i++) r[i * 4] = 1;
  return r;
}

function decodeShSidecar(manifest, count, shBytes) {
  if (manifest.sh === undefined) return undefined;
  var sh = requireRecord(manifest.sh, 'sh');
  var degree = requireNonNegativeInteger(sh.degree, 'sh.degree');
  if (degree <= 0) throw new Error('sh.degree must be greater than zero when an SH sidecar is declared');
  if (sh.basis !== '3dgs_real_sh') throw new Error('sh.basis must be 3dgs_real_sh');
  if (sh.coefficients_layout !== 'splat_coeff_rgb') throw new Error('sh.coefficients_layout must be splat_coeff_rgb');
  if (sh.coefficients_component_type !== 'float32') throw new Error('sh.coefficients_component_type must be float32');
  var k = (degree + 1) * (degree + 1) - 1;
  var declared = requirePositiveInteger(sh.coefficient_count, 'sh.coefficient_count');
  if (declared !== k) throw new Error('sh.coefficient_count must be ' + k + ' for degree ' + degree);
  var expected = count * k * 3 * 4;
  var declaredBytes = requirePositiveInteger(sh.coefficients_byte_length, 'sh.coefficients_byte_length');
  if (declaredBytes !== expected) {
    throw new Error('sh.coefficients_byte_length must be ' + expected + ' for ' + count + ' splats at degree ' + degree);
  }
  if (shBytes === undefined) throw new Error('SH sidecar bytes are required when manifest.sh is declared');
  return { degree: degree, basis: '3dgs_real_sh', coefficientCount: k, layout: 'splat_coeff_rgb',
    coefficients: decodeFloat32Bytes(shBytes, count * k * 3, 'sh.coefficients') };
}

function manifestSourceKind(m) {
  if (m.source !== undefined) {
    var src = requireRecord(m.source, 'source');
    if (src.kind !== undefined) {
      var kind = requireString(src.kind, 'source.kind');
      return kind === 'spz' ? 'real_splat_spz' : kind;
    }
  }
  return 'real_scaniverse_ply';
}

function decodeFirstSmokeSplatManifest(manifest, payloadBytes, idsBytes, scalesBytes, rotationsBytes, shBytes,
                                       roughnessBytes, metalnessBytes) {
  var root = requireRecord(manifest, 'manifest');
  var count = requirePositiveInteger(firstDefined(root.splat_count, root.count), 'splat_count');
  if (firstDefined(root.endianness, 'little') !== 'little') throw new Error('manifest.endianness must be little');
  var bounds = decodeBounds(root.bounds, 'bounds');
  var sourceKind = manifestSourceKind(root);
  var layout = decodeLayout(layoutSource(root), 'layout');
  var info = requireRecord(root.payload, 'payload');
  var expected = count * layout.strideBytes;
  var declared = requirePositiveInteger(firstDefined(info.byte_length, expected), 'payload.byte_length');
  if (declared !== expected) throw new Error('payload.byte_length must be ' + expected + ' for ' + count + ' rows');
  var ab = asArrayBuffer(payloadBytes);
  if (ab.byteLength !== expected) {
    throw new Error('payload sidecar byte length ' + ab.byteLength + ' does not match expected ' + expected);
  }
  var data = new DataView(ab);
  var positions = new Float32Array(count * 3), colors = new Float32Array(count * 3);
  var opacities = new Float32Array(count), radii = new Float32Array(count);
  for (var row = 0; row < count; row++) {
    var base = row * 32;
    for (var c = 0; c < 3; c++) {
      positions[row * 3 + c] = data.getFloat32(base + c * 4, true);
      colors[row * 3 + c] = data.getFloat32(base + 12 + c * 4, true);
    }
    opacities[row] = requireUnitInterval(data.getFloat32(base + 24, true), 'payload.rows[' + row + '].opacity');
    radii[row] = requirePositiveFinite(data.getFloat32(base + 28, true), 'payload.rows[' + row + '].radius');
  }
  var ids;
  if (idsBytes === undefined) {
    ids = decodeOriginalIds(undefined, count);
  }
  var result = {
    count: count,
    sourceKind: sourceKind,
    positions: positions,
    colors: colors,
    opacities: opacities,
    radii: radii,
    scales: scalesBytes === undefined ? fallbackScalesFromRadii(radii) : decodeFloat32Bytes(scalesBytes, count * 3, 'shape.scales'),
    rotations: rotationsBytes === undefined ? identityRotations(count) : decodeFloat32Bytes(rotationsBytes, count * 4, 'shape.rotations'),
    sh: decodeShSidecar(root, count, shBytes),
    roughness: roughnessBytes === undefined ? undefined : decodeFloat32Bytes(roughnessBytes, count, 'materials.roughness'),
    metalness: metalnessBytes === undefined ? undefined : decodeFloat32Bytes(metalnessBytes, count, 'materials.metalness'),
    originalIds: ids,
    bounds: bounds,
    layout: layout,
  };
  if (idsBytes !== undefined) {
    var ib = asArrayBuffer(idsBytes);
    if (ib.byteLength !== count * 4) throw new Error('ID sidecar byte length ' + ib.byteLength + ' does not match expected ' + count * 4);
    var dv = new DataView(ib);
    var raw = new Uint32Array(count);
    for (var i = 0; i < count; i++) raw[i] = dv.getUint32(i * 4, true);
    result.originalIds = decodeOriginalIds(raw, count);
  }
  return result;
}

function decodeFloat32Array(value, expectedLength, p, validator) {
  validator = validator || requireFiniteNumber;
  if (!Array.isArray(value) && !ArrayBuffer.isView(value)) throw new Error(p + ' must be an array or typed array');
  var values = Array.prototype.slice.call(value);
  if (values.length !== expectedLength) throw new Error(p + ' must contain ' + expectedLength + ' entries');
  var out = new Float32Array(expectedLength);
  for (var i = 0; i < values.length; i++) out[i] = validator(values[i], p + '[' + i + ']');
  return out;
}

function decodeFirstSmokeSplatPayload(payload) {
  var root = requireRecord(payload, 'payload');
  var meta = root.metadata === undefined ? root : requireRecord(root.metadata, 'metadata');
  var count = requirePositiveInteger(firstDefined(meta.count, meta.splat_count), 'count');
  var bounds = decodeBounds(meta.bounds, 'bounds');
  var layout = decodeLayout(layoutSource(meta), 'layout');
  var d;
  if (root.splats !== undefined) {
    if (!Array.isArray(root.splats)) throw new Error('splats must be an array of first-smoke splat rows');
    if (root.splats.length !== count) throw new Error('count ' + count + ' does not match ' + root.splats.length + ' rows');
    d = { positions: new Float32Array(count * 3), colors: new Float32Array(count * 3), opacities: new Float32Array(count),
      radii: new Float32Array(count) };
    root.splats.forEach(function (r, i) {
      var row = requireRecord(r, 'splats[' + i + ']');
      d.positions.set(requireVec3(row.position, 'splats[' + i + '].position'), i * 3);
      d.colors.set(requireVec3(row.color, 'splats[' + i + '].color'), i * 3);
      d.opacities[i] = requireUnitInterval(row.opacity, 'splats[' + i + '].opacity');
      d.radii[i] = requirePositiveFinite(row.radius, 'splats[' + i + '].radius');
    });
    d.scales = fallbackScalesFromRadii(d.radii);
    d.rotations = identityRotations(count);
  } else {
    var a = requireRecord(root.attributes, 'attributes');
    d = {
      positions: decodeFloat32Array(firstDefined(a.positions, a.position), count * 3, 'attributes.positions'),
      colors: decodeFloat32Array(firstDefined(a.colors, a.color), count * 3, 'attributes.colors'),
      opacities: decodeFloat32Array(firstDefined(a.opacities, a.opacity), count, 'attributes.opacities', requireUnitInterval),
      radii: decodeFloat32Array(firstDefined(a.radii, a.radius), count, 'attributes.radii', requirePositiveFinite),
    };
    d.scales = a.scales === undefined ? fallbackScalesFromRadii(d.radii) : decodeFloat32Array(a.scales, count * 3, 'attributes.scales');
    d.rotations = a.rotations === undefined ? identityRotations(count) : decodeFloat32Array(a.rotations, count * 4, 'attributes.rotations');
  }
  return {
    count: count,
    sourceKind: requireString(firstDefined(meta.sourceKind, 'real_scaniverse_ply'), 'sourceKind'),
    positions: d.positions, colors: d.colors, opacities: d.opacities, radii: d.radii, scales: d.scales,
    rotations: d.rotations, originalIds: decodeOriginalIds(root.originalIds, count), bounds: bounds, layout: layout,
  };
}

function isSidecarManifest(p) {
  return !!p && typeof p === 'object' && !Array.isArray(p) && !!p.payload && typeof p.payload === 'object' &&
    !Array.isArray(p.payload) && typeof p.payload.path === 'string';
}

function resolveSidecar(manifestPath, p) {
  return path.isAbsolute(p) ? p : path.join(path.dirname(manifestPath), p);
}

function loadFirstSmokeSplatPayload(file) {
  var payload = JSON.parse(fs.readFileSync(file, 'utf8'));
  if (!isSidecarManifest(payload)) return decodeFirstSmokeSplatPayload(payload);
  var root = requireRecord(payload, 'payload');
  var info = requireRecord(root.payload, 'payload.payload');
  var read = function (p) { return p === undefined ? undefined : fs.readFileSync(resolveSidecar(file, p)); };
  var idsPath;
  if (root.ids !== undefined && requireRecord(root.ids, 'ids').path !== undefined) idsPath = requireString(root.ids.path, 'ids.path');
  else if (root.identity !== undefined && requireRecord(root.identity, 'identity').ids_path !== undefined) {
    idsPath = requireString(root.identity.ids_path, 'identity.ids_path');
  }
  var shape = root.shape === undefined ? {} : requireRecord(root.shape, 'shape');
  var scalesPath = shape.scales_path === undefined ? undefined : requireString(shape.scales_path, 'shape.scales_path');
  var rotPath = shape.rotations_path === undefined ? undefined : requireString(shape.rotations_path, 'shape.rotations_path');
  var shPath = root.sh === undefined ? undefined : requireString(requireRecord(root.sh, 'sh').coefficients_path, 'sh.coefficients_path');
  var mats = root.materials || {};
  return decodeFirstSmokeSplatManifest(payload, read(requireString(info.path, 'payload.path')), read(idsPath),
    read(scalesPath), read(rotPath), read(shPath), read(mats.roughness_path), read(mats.metalness_path));
}

function fetchFirstSmokeSplatPayload(file) {
  return new Promise(function (resolve) { resolve(loadFirstSmokeSplatPayload(file)); });
}

// ---- packers (src/splats.ts:431-500) ----
function f32ToF16(value) {  // the host's truncating conversion (src/splats.ts:431-442)
  var f = new Float32Array(1), u = new Uint32Array(f.buffer);
  f[0] = value;
  var bits = u[0];
  var sign = (bits >>> 16) & 0x8000;
  var exp = ((bits >>> 23) & 0xFF) - 127 + 15;
  var frac = bits & 0x7FFFFF;
  if (exp <= 0) return sign;
  if (exp >= 31) return sign | 0x7C00;
  return sign | (exp << 10) | (frac >>> 13);
}

function packMaterialBuffer(count, roughness, metalness) {
  var out = new Uint32Array(count);
  for (var i = 0; i < count; i++) {
    var r = roughness ? roughness[i] : 0.75, m = metalness ? metalness[i] : 0.0;
    out[i] = (f32ToF16(r) | (f32ToF16(m) << 16)) >>> 0;
  }
  return out;
}

function packNormalBuffer(count, normals, detailNormals) {
  if (!detailNormals) return normals;
  var out = new Float32Array(count * 6);
  out.set(normals, 0);
  out.set(detailNormals, count * 3);
  return out;
}

function packShDataBuffer(dcColors, sh, emissive) {
  var count = dcColors.length / 3;
  var k = sh ? sh.coefficientCount : 0;
  var stride = 3 + k * 3 + 3;
  var out = new Float32Array(count * stride);
  for (var i = 0; i < count; i++) {
    var o = i * stride;
    out[o] = dcColors[i * 3]; out[o + 1] = dcColors[i * 3 + 1]; out[o + 2] = dcColors[i * 3 + 2];
    if (sh) out.set(sh.coefficients.subarray(i * k * 3, (i + 1) * k * 3), o + 3);
    if (emissive) {
      out[o + 3 + k * 3] = emissive[i * 3]; out[o + 4 + k * 3] = emissive[i * 3 + 1]; out[o + 5 + k * 3] = emissive[i * 3 + 2];
    }
  }
  return out;
}

function uploadSplatAttributeBuffers(device, attributes) {
  var a = attributes;
  return {
    count: a.count,
    positionBuffer: a.positions,
    opacityBuffer: a.opacities,
    scaleBuffer: a.scales,
    rotationBuffer: a.rotations,
    materialBuffer: packMaterialBuffer(a.count, a.roughness, a.metalness),
    normalBuffer: a.normals ? packNormalBuffer(a.count, a.normals, a.detailNormals) : undefined,
    shDataBuffer: packShDataBuffer(a.colors, a.sh, a.emissive),
    shDegree: a.sh ? a.sh.degree : 0,
    originalIdBuffer: a.originalIds,
    bounds: a.bounds,
    layout: a.layout,
  };
}

function framingFromBounds(bounds, options) {
  var o = options || {};
  var fovY = o.fovY === undefined ? Math.PI / 3 : o.fovY;
  var padding = o.padding === undefined ? 1.35 : o.padding;
  var minNear = o.minNear === undefined ? 0.001 : o.minNear;
  if (!isFinite(fovY) || fovY <= 0 || fovY >= Math.PI) throw new Error('framing.fovY must be a finite angle in (0, pi)');
  if (!isFinite(padding) || padding < 1) throw new Error('framing.padding must be a finite value >= 1');
  var dx = bounds.max[0] - bounds.min[0], dy = bounds.max[1] - bounds.min[1], dz = bounds.max[2] - bounds.min[2];
  var radius = Math.max(bounds.radius, Math.hypot(dx, dy, dz) / 2);
  var padded = radius * padding;
  var distance = padded / Math.sin(fovY / 2);
  return { target: bounds.center.slice(), radius: radius, distance: distance,
    near: Math.max(minNear, distance - padded * 2), far: distance + padded * 2 };
}

module.exports = {
  FIRST_SMOKE_SPLAT_LAYOUT: FIRST_SMOKE_SPLAT_LAYOUT,
  decodeLocalPlySplatPayload: decodeLocalPlySplatPayload,
  decodeFirstSmokeSplatManifest: decodeFirstSmokeSplatManifest,
  decodeFirstSmokeSplatPayload: decodeFirstSmokeSplatPayload,
  loadFirstSmokeSplatPayload: loadFirstSmokeSplatPayload,
  fetchFirstSmokeSplatPayload: fetchFirstSmokeSplatPayload,
  uploadSplatAttributeBuffers: uploadSplatAttributeBuffers,
  framingFromBounds: framingFromBounds,
};
