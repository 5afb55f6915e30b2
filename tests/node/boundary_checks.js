'use strict';
// Error behaviour and ownership rules of the JS boundary, on a scene written by the Python test:
//  * every splat buffer shorter than splatCount requires, and a short sortedIndices update, throw
//    a RangeError before anything is read; a shData length that matches no SH degree throws;
//  * frame uniforms whose shDegree differs from the bound shData's are rejected;
//  * while a renderAsync() frame is in flight every other call throws "busy", and destroy() is
//    deferred to the frame's completion (the Promise still resolves; the handle is dead after);
//  * an output texture given as a device address (createTexture({devicePointer})) receives the
//    frame: another context's own output buffers are bound and must hold the image that context
//    renders itself.
// Prints one JSON line with what happened.
var fs = require('fs');
var path = require('path');
var pkg = process.argv[2], dir = process.argv[3];
var C = require(path.join(pkg, 'node', 'gpuTileSplatCompositor.js'));
var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function f32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }
function attempt(fn) {
  try { fn(); return null; } catch (e) { return { name: e.name, message: String(e.message) }; }
}
var device = C.createDevice();
var plan = C.planTileSplatCompositor({ viewportWidth: meta.width, viewportHeight: meta.height, tileSizePx: 16, splatCount: meta.count });
var buffers = {
  positionBuffer: f32('positions.bin'), scaleBuffer: f32('scales.bin'), rotationBuffer: f32('rotations.bin'),
  opacityBuffer: f32('opacities.bin'), materialBuffer: u32('materials.bin'), shDataBuffer: f32('shdata.bin'),
  sortedIndexBuffer: u32('sorted.bin'),
};
var uniforms = new Float32Array(64);
C.writeTileSplatFrameUniforms(uniforms, f32('viewproj.bin'), plan, 1.0, meta.shDegree, f32('campos.bin'), f32('view.bin'), f32('proj.bin'));
var out = { short: {} };
var res = C.createTileSplatCompositor(device, plan, {});
Object.keys(buffers).forEach(function (k) {
  var b = {};
  Object.keys(buffers).forEach(function (j) { b[j] = buffers[j]; });
  b[k] = buffers[k].subarray(0, buffers[k].length - 1);
  out.short[k] = attempt(function () { C.createTileSplatBindGroups(device, res, b, null, null); });
});
var odd = {};
Object.keys(buffers).forEach(function (j) { odd[j] = buffers[j]; });
odd.shDataBuffer = new Float32Array(buffers.shDataBuffer.length + meta.count);  // 6+3K+1 floats per splat
out.noDegree = attempt(function () { C.createTileSplatBindGroups(device, res, odd, null, null); });
var bg = C.createTileSplatBindGroups(device, res, buffers, null, null);
out.shortSorted = attempt(function () { res.updateSortedIndices(buffers.sortedIndexBuffer.subarray(1)); });
var wrong = new Float32Array(uniforms);
new Uint32Array(wrong.buffer)[25] = meta.shDegree === 0 ? 3 : 0;
out.wrongDegree = attempt(function () { device.queue.writeBuffer(res.frameUniformBuffer, 0, wrong); });
device.queue.writeBuffer(res.frameUniformBuffer, 0, uniforms);
C.encodeFullComputeCompositorPipeline(device.createCommandEncoder(), res, bg);
device.queue.submit([bg]);
var image = res.readRGBA();

// an in-flight renderAsync frame: other calls throw, destroy() waits for it
var p = res.renderAsync();
out.busyRender = attempt(function () { C.encodeFullComputeCompositorPipeline(null, res, bg); });
out.busyRead = attempt(function () { res.readRGBA(); });
out.busySecondAsync = attempt(function () { res.renderAsync(); });
out.destroyDuringFlight = attempt(function () { res.destroy(); });
p.then(function () {
  out.asyncResolved = true;
  out.afterDestroy = attempt(function () { res.readRGBA(); });
  // caller-owned device outputs: a second context renders into the first one's buffers
  var owner = C.createTileSplatCompositor(device, plan, {});
  var addr = owner.deviceOutputs();
  var res2 = C.createTileSplatCompositor(device, plan, {});
  var colour = device.createTexture({ size: [meta.width, meta.height], format: 'rgba32float', devicePointer: addr.rgba });
  var depth = device.createTexture({ size: [meta.width, meta.height], format: 'r32float', devicePointer: addr.depth });
  var bg2 = C.createTileSplatBindGroups(device, res2, buffers, colour, { depth: depth });
  device.queue.writeBuffer(res2.frameUniformBuffer, 0, uniforms);
  C.encodeFullComputeCompositorPipeline(null, res2, bg2);
  device.queue.submit([bg2]);
  var got = owner.readRGBA();
  var equal = got.length === image.length;
  for (var i = 0; equal && i < got.length; i++) if (got[i] !== image[i]) equal = false;
  out.deviceOutputEqual = equal;
  out.deviceOutputNonzero = addr.rgba !== BigInt(0);
  res2.destroy();
  owner.destroy();
  console.log(JSON.stringify(out));
}, function (e) {
  console.error('renderAsync rejected: ' + e);
  process.exitCode = 1;
});
