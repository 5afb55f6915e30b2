'use strict';
// gpuTileSplatCompositor.js - drop-in facade of the reference's src/gpuTileSplatCompositor.ts
// over the MI355X compositor (hdsr_napi.node -> libhdsr.so).  Node >= 12, CommonJS, no optional
// chaining or nullish coalescing.
//
// Exports the reference's names with the same argument meaning:
//   TILE_SPLAT_FRAME_UNIFORM_BYTES                                   (:18)
//   planTileSplatCompositor(input)                                   (:109-139)
//   createTileSplatCompositor(device, plan, options)                 (:141-728)
//   writeTileSplatFrameUniforms(target, viewProj, plan, splatScale, shDegree, cameraPos,
//                               viewMatrix, projMatrix, normalMatrix) (:730-774)
//   createTileSplatBindGroups(device, resources, splatBuffers, outputTexture, gbufferTextures)
//                                                                    (:792-995)
//   encodeFullComputeCompositorPipeline(encoder, resources, bindGroups) (:1023-1179)
//   encodeCompositeOnly(encoder, resources, bindGroups)              (:1001-1014)
// GPU objects become typed arrays: splat buffers are Float32Array / Uint32Array in the layouts of
// src/splats.ts:376-500 (bound with the reference's keys; the SH degree comes from the shData
// length, as the reference takes it from the frame uniforms), the frame uniform buffer is written
// with device.queue.writeBuffer(resources.frameUniformBuffer, 0, float32x64) as the reference's
// caller does (src/splatRenderer.ts:1030).  The output textures are what
// device.createTexture({size: [W, H], format}) returns: host textures whose `data` typed array
// device.queue.submit() fills from the finished frame, or, with `devicePointer` (a BigInt device
// address), caller-owned device buffers the frame is written into (rgba32float colour, r32float
// depth).  resources.readRGBA() / readDepth() read the context's own outputs.  createDevice()
// returns the minimal device object these calls need.

var path = require('path');
var addon = require(path.join(__dirname, 'hdsr_napi.node'));

var TILE_SPLAT_FRAME_UNIFORM_BYTES = 256;

var FLAGS = { f16: 1, refCaps: 2, gbuffer: 4, debug: 8, rgba16: 16, profile: 32, bucketScatter: 128 };
// texture format -> (output buffer, typed array constructor, elements per pixel, creation flag it needs)
var TEXTURE_FORMATS = {
  rgba32float: { out: 0, ctor: Float32Array, per: 4, flag: 0 },
  rgba16float: { out: 4, ctor: Uint16Array, per: 4, flag: 0 },
  r32float: { out: 1, ctor: Float32Array, per: 1, flag: 0 },
  r32uint: { out: 2, ctor: Uint32Array, per: 1, flag: 4 },
  rgba32uint: { out: 3, ctor: Uint32Array, per: 4, flag: 4 },
};
var OUT = { rgba: 0, depth: 1, gbufferNormal: 2, gbufferMaterial: 3, rgba16: 4, filteredNormal: 7, lit: 8 };
var DBG = {
  projCache: 0, depth: 1, tileCounts: 2, tileOffsets: 3, entries: 4, entriesCanon: 5, smallTileList: 6,
  largeTileList: 7, largeTileOverflowBases: 8, chunkRanges: 9, keys20: 10, largeSplatList: 11, sortedIndices: 12,
};

function planTileSplatCompositor(input) {
  return addon.planCompute({
    viewportWidth: input.viewportWidth,
    viewportHeight: input.viewportHeight,
    tileSizePx: input.tileSizePx,
    splatCount: input.splatCount,
    tileEntryMultiplier: input.tileEntryMultiplier === undefined ? 0 : input.tileEntryMultiplier,
  });
}

function FrameUniformBuffer(resources) {
  this.label = 'tile_splat_frame_uniforms';
  this.size = TILE_SPLAT_FRAME_UNIFORM_BYTES;
  this._resources = resources;
}
FrameUniformBuffer.prototype._write = function (data) {
  var view = data instanceof Float32Array ? data :
    new Float32Array(data.buffer, data.byteOffset, TILE_SPLAT_FRAME_UNIFORM_BYTES / 4);
  addon.setFrame(this._resources.handle, view);
};

function createTileSplatCompositor(device, plan, options) {
  var opts = options || {};
  var flags = 0;
  Object.keys(FLAGS).forEach(function (k) { if (opts[k]) flags |= FLAGS[k]; });
  var handle = addon.create(plan, flags);
  var px = plan.viewportWidth * plan.viewportHeight;
  var resources = {
    plan: plan,
    handle: handle,
    flags: flags,
    destroy: function () { addon.destroy(handle); },
    sync: function () { addon.sync(handle); },
    // the full pipeline on a worker thread: a Promise settled when the frame is complete (the
    // context must not be used until then)
    renderAsync: function () { return addon.renderAsync(handle); },
    readRGBA: function () { var a = new Float32Array(px * 4); addon.readback(handle, OUT.rgba, a); return a; },
    readDepth: function () { var a = new Float32Array(px); addon.readback(handle, OUT.depth, a); return a; },
    readOutput: function (name, array) { addon.readback(handle, OUT[name], array); return array; },
    debug: function (name) { return addon.debugReadback(handle, DBG[name]); },
    stats: function () { return addon.stats(handle); },
    stageTimes: function () { return addon.stageTimes(handle); },
    updateSortedIndices: function (sorted) { addon.updateSortedIndices(handle, sorted, plan.splatCount); },
    deviceOutputs: function () { return addon.outputs(handle); },
    _targets: [],
  };
  resources.frameUniformBuffer = new FrameUniformBuffer(resources);
  return resources;
}

// Byte-identical 256-byte layout of the reference (gpu_project_splats.wgsl:27-44).
function writeTileSplatFrameUniforms(target, viewProj, plan, splatScale, shDegree, cameraPos, viewMatrix,
                                     projMatrix, normalMatrix) {
  if (splatScale === undefined) splatScale = 1.0;
  if (shDegree === undefined) shDegree = 0;
  target.set(viewProj, 0);
  target[16] = plan.viewportWidth;
  target[17] = plan.viewportHeight;
  target[18] = plan.tileSizePx;
  target[19] = 0;
  var u32 = new Uint32Array(target.buffer, target.byteOffset, target.length);
  u32[20] = plan.tileColumns;
  u32[21] = plan.tileRows;
  u32[22] = plan.splatCount;
  u32[23] = plan.maxTotalTileRefs;
  target[24] = splatScale;
  u32[25] = shDegree;
  if (cameraPos) {
    target[28] = cameraPos[0];
    target[29] = cameraPos[1];
    target[30] = cameraPos[2];
  }
  if (viewMatrix) target.set(viewMatrix, 32);
  if (projMatrix) {
    target[48] = plan.viewportWidth * projMatrix[0] * 0.5;
    target[49] = plan.viewportHeight * projMatrix[5] * 0.5;
  }
  var n = normalMatrix || [1, 0, 0, 0, 1, 0, 0, 0, 1];
  target[52] = n[0]; target[53] = n[1]; target[54] = n[2]; target[55] = 0;
  target[56] = n[3]; target[57] = n[4]; target[58] = n[5]; target[59] = 0;
  target[60] = n[6]; target[61] = n[7]; target[62] = n[8]; target[63] = 0;
}

// Checks a texture bound as output 'role' and returns its target: {out, texture} for a host
// texture (filled at queue.submit) or {device: BigInt} for a device buffer.
function outputTarget(resources, texture, role, formats) {
  var plan = resources.plan;
  var f = TEXTURE_FORMATS[texture.format];
  if (!f || formats.indexOf(texture.format) < 0) {
    throw new TypeError(role + ' texture format must be one of ' + formats.join(', ') + ' (got ' + texture.format + ')');
  }
  if (texture.width !== plan.viewportWidth || texture.height !== plan.viewportHeight) {
    throw new RangeError(role + ' texture is ' + texture.width + 'x' + texture.height + '; the plan is ' +
                         plan.viewportWidth + 'x' + plan.viewportHeight);
  }
  if (f.flag && !(resources.flags & f.flag)) {
    throw new Error(role + ' texture format ' + texture.format + ' needs the compositor created with {gbuffer: true}');
  }
  if (texture.devicePointer !== undefined) {
    if (texture.format !== 'rgba32float' && texture.format !== 'r32float') {
      throw new TypeError(role + ': only rgba32float colour and r32float depth can be device buffers');
    }
    return { device: texture.devicePointer };
  }
  if (texture.format === 'rgba16float' && !(resources.flags & FLAGS.rgba16)) {
    return { out: OUT.rgba, texture: texture, toHalf: true };  // fp32 read back, rounded here
  }
  return { out: f.out, texture: texture };
}

// IEEE binary16 bits of each float32 value, round to nearest even (subnormals, overflow to
// infinity, NaN kept quiet): what the rgba16float texture stores for the fp32 colour.
var f32buf = new Float32Array(1), u32buf = new Uint32Array(f32buf.buffer);
function halfBits(x) {
  f32buf[0] = x;
  var b = u32buf[0];
  var sign = (b >>> 16) & 0x8000, e = (b >>> 23) & 0xFF, m = b & 0x7FFFFF;
  if (e === 0xFF) return sign | 0x7C00 | (m ? 0x200 | (m >>> 13) : 0);
  var he = e - 127 + 15;
  if (he >= 0x1F) return sign | 0x7C00;
  if (he <= 0) {
    if (he < -10) return sign;
    m |= 0x800000;
    var shift = 14 - he;
    var h = m >>> shift, rem = m & ((1 << shift) - 1), half = 1 << (shift - 1);
    if (rem > half || (rem === half && (h & 1))) h++;
    return sign | h;
  }
  var hm = m >>> 13, r = m & 0x1FFF;
  var out = sign | (he << 10) | hm;
  if (r > 0x1000 || (r === 0x1000 && (hm & 1))) out++;
  return out;
}

function toHalfArray(src, dst) {
  for (var i = 0; i < src.length; i++) dst[i] = halfBits(src[i]);
}

function createTileSplatBindGroups(device, resources, splatBuffers, outputTexture, gbufferTextures) {
  var buffers = {
    positions: splatBuffers.positionBuffer,
    scales: splatBuffers.scaleBuffer,
    rotations: splatBuffers.rotationBuffer,
    opacities: splatBuffers.opacityBuffer,
    materials: splatBuffers.materialBuffer,
    normals: splatBuffers.normalBuffer,
    shData: splatBuffers.shDataBuffer,
    sortedIndices: splatBuffers.sortedIndexBuffer,
  };
  // the reference's bind names no SH degree: without one the addon takes it from shData's length
  if (splatBuffers.shDegree !== undefined) buffers.shDegree = splatBuffers.shDegree;
  addon.bindSplats(resources.handle, resources.plan.splatCount, buffers);
  var targets = [];
  var rgbaDev = null, depthDev = null;
  if (outputTexture) {
    var t = outputTarget(resources, outputTexture, 'outputTexture', ['rgba32float', 'rgba16float']);
    if (t.device !== undefined) rgbaDev = t.device; else targets.push(t);
  }
  var gb = gbufferTextures || {};
  if (gb.depth) {
    var d = outputTarget(resources, gb.depth, 'gbufferTextures.depth', ['r32float']);
    if (d.device !== undefined) depthDev = d.device; else targets.push(d);
  }
  if (gb.normal) targets.push(outputTarget(resources, gb.normal, 'gbufferTextures.normal', ['r32uint']));
  if (gb.material) targets.push(outputTarget(resources, gb.material, 'gbufferTextures.material', ['rgba32uint']));
  addon.bindOutputs(resources.handle, rgbaDev, depthDev);
  resources._targets = targets;
  return { resources: resources, outputTexture: outputTexture, gbufferTextures: gbufferTextures };
}

// Copies the finished frame into the host textures bound by createTileSplatBindGroups.
function fillTargets(resources) {
  resources._targets.forEach(function (t) {
    if (t.toHalf) {
      var tmp = new Float32Array(t.texture.data.length);
      addon.readback(resources.handle, t.out, tmp);
      toHalfArray(tmp, t.texture.data);
    } else {
      addon.readback(resources.handle, t.out, t.texture.data);
    }
  });
}

function encodeFullComputeCompositorPipeline(encoder, resources, bindGroups) {
  addon.render(resources.handle);
}

// The GPU sortedIndices producer (createViewDepthSortInput + encodeGpuSortPrototype,
// src/gpuSortPrototype.ts:120-140, 275-287) for the current frame uniforms' view matrix.
function encodeViewDepthSort(encoder, resources) {
  addon.sortBackToFront(resources.handle);
}

// The first deferred consumers of the G-buffer (src/splatRenderer.ts:1042-1069): screen-space
// normal reconstruction into the filtered-normal buffer, then the bilateral filter back into the
// G-buffer normal words.  options: recoveryMode (0 screen-space, 1 grazing recovery, 2 source
// pass-through), grazingStart / grazingEnd, spatialSigma / depthSigma / normalSigma /
// kernelRadius; unset fields take the reference's values.
function encodeScreenSpaceNormals(encoder, resources, options) {
  addon.screenSpaceNormals(resources.handle, options || {});
}

function encodeBilateralNormalFilter(encoder, resources, options) {
  addon.bilateralNormalFilter(resources.handle, options || {});
}

// Deferred lighting (createDeferredLightingPass().encode, src/splatRenderer.ts:446-548, as encoded
// at :1115-1139): colour, depth, G-buffer normals (after the two passes above) and materials ->
// the lit rgba16float buffer (readback output 'lit').  options take encode's arguments by name:
// lightDirection [x, y, z], lightIntensity, ambientIntensity, specularOnly, emissiveIntensity,
// emissiveThreshold, envIntensity, envRotation, exposure; the MaterialCurveParams roughnessCurve /
// metalnessCurve / albedoCurve ({contrast, brightness, gamma}); lightingViewProj (16 numbers) and
// lightingCameraPosition (3) for a separate lighting camera; ao and bloom as device addresses
// (BigInt) of the AO texture (H*W*4 f32) and the half-resolution bloom texture, or null.  Unset
// fields take the reference viewer's values; the environment-map path is not built.
function encodeDeferredLighting(encoder, resources, options) {
  addon.deferredLighting(resources.handle, options || {});
}

function encodeCompositeOnly(encoder, resources, bindGroups) {
  addon.compositeOnly(resources.handle);
}

// The device object the calls above need: queue.writeBuffer for the frame uniforms, a
// queue.submit that waits for the enqueued work (and fills the bound host textures), and
// createTexture for the output textures (host storage; {devicePointer} makes a device buffer).
function createDevice() {
  var pending = [];
  return {
    createTexture: function (desc) {
      var size = desc.size || [];
      var w = size[0] !== undefined ? size[0] : size.width, h = size[1] !== undefined ? size[1] : size.height;
      var f = TEXTURE_FORMATS[desc.format];
      if (!f) throw new TypeError('unsupported texture format ' + desc.format);
      var tex = { label: desc.label, width: w, height: h, format: desc.format };
      if (desc.devicePointer !== undefined) tex.devicePointer = desc.devicePointer;
      else tex.data = new f.ctor(w * h * f.per);
      return tex;
    },
    queue: {
      writeBuffer: function (buffer, offset, data) {
        if (!(buffer instanceof FrameUniformBuffer) || offset !== 0) {
          throw new Error('only the compositor frame uniform buffer can be written');
        }
        buffer._write(data);
      },
      submit: function (commandBuffers) {
        (commandBuffers || []).forEach(function (cb) { if (cb && cb.resources) pending.push(cb.resources); });
        while (pending.length) {
          var r = pending.pop();
          r.sync();
          fillTargets(r);
        }
      },
    },
    createCommandEncoder: function () {
      return { finish: function () { return {}; } };
    },
  };
}

module.exports = {
  TILE_SPLAT_FRAME_UNIFORM_BYTES: TILE_SPLAT_FRAME_UNIFORM_BYTES,
  planTileSplatCompositor: planTileSplatCompositor,
  createTileSplatCompositor: createTileSplatCompositor,
  writeTileSplatFrameUniforms: writeTileSplatFrameUniforms,
  createTileSplatBindGroups: createTileSplatBindGroups,
  encodeFullComputeCompositorPipeline: encodeFullComputeCompositorPipeline,
  encodeCompositeOnly: encodeCompositeOnly,
  encodeViewDepthSort: encodeViewDepthSort,
  encodeScreenSpaceNormals: encodeScreenSpaceNormals,
  encodeBilateralNormalFilter: encodeBilateralNormalFilter,
  encodeDeferredLighting: encodeDeferredLighting,
  createDevice: createDevice,
  version: addon.version,
};
