'use strict';
// Renders one frame through resources.renderAsync() (the pipeline on a worker thread) from the
// binaries written by the Python test, counts the event-loop turns that ran while the Promise was
// pending, then renders the same frame synchronously and reports whether the two images agree.
var fs = require('fs');
var path = require('path');
var pkg = process.argv[2], dir = process.argv[3];
var C = require(path.join(pkg, 'node', 'gpuTileSplatCompositor.js'));
var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function f32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }
var device = C.createDevice();
var plan = C.planTileSplatCompositor({ viewportWidth: meta.width, viewportHeight: meta.height, tileSizePx: 16, splatCount: meta.count });
var resources = C.createTileSplatCompositor(device, plan, {});
var bindGroups = C.createTileSplatBindGroups(device, resources, {
  positionBuffer: f32('positions.bin'), scaleBuffer: f32('scales.bin'), rotationBuffer: f32('rotations.bin'),
  opacityBuffer: f32('opacities.bin'), materialBuffer: u32('materials.bin'), shDataBuffer: f32('shdata.bin'),
  sortedIndexBuffer: u32('sorted.bin'),
}, null, null);
var target = new Float32Array(C.TILE_SPLAT_FRAME_UNIFORM_BYTES / 4);
C.writeTileSplatFrameUniforms(target, f32('viewproj.bin'), plan, 1.0, meta.shDegree, f32('campos.bin'), f32('view.bin'), f32('proj.bin'));
device.queue.writeBuffer(resources.frameUniformBuffer, 0, target);
var ticks = 0, pending = true;
function tick() { if (pending) { ticks++; setImmediate(tick); } }
setImmediate(tick);
resources.renderAsync().then(function () {
  pending = false;
  var a = resources.readRGBA();
  fs.writeFileSync(path.join(dir, 'rgba_async.bin'), Buffer.from(a.buffer));
  C.encodeFullComputeCompositorPipeline(device.createCommandEncoder(), resources, bindGroups);
  device.queue.submit([bindGroups]);
  var b = resources.readRGBA();
  var equal = a.length === b.length;
  for (var i = 0; equal && i < a.length; i++) if (a[i] !== b[i] && !(a[i] !== a[i] && b[i] !== b[i])) equal = false;
  console.log(JSON.stringify({ equal: equal, ticks: ticks, stats: resources.stats() }));
  resources.destroy();
}, function (e) {
  pending = false;
  console.error(String(e));
  process.exitCode = 1;
});
