'use strict';
// Renders one frame through the JS facade in the reference's call order
// (src/splatRenderer.ts:901-961 load, :1029-1040 frame) from binaries written by the Python test,
// and writes the outputs and intermediates back next to them.
var fs = require('fs');
var path = require('path');
var pkg = process.argv[2], dir = process.argv[3];
var C = require(path.join(pkg, 'node', 'gpuTileSplatCompositor.js'));
var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function f32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }
var device = C.createDevice();
var plan = C.planTileSplatCompositor({ viewportWidth: meta.width, viewportHeight: meta.height, tileSizePx: 16, splatCount: meta.count });
var resources = C.createTileSplatCompositor(device, plan, { debug: true, gbuffer: !!meta.gbuffer });
var bindGroups = C.createTileSplatBindGroups(device, resources, {
  positionBuffer: f32('positions.bin'), scaleBuffer: f32('scales.bin'), rotationBuffer: f32('rotations.bin'),
  opacityBuffer: f32('opacities.bin'), materialBuffer: u32('materials.bin'), shDataBuffer: f32('shdata.bin'),
  sortedIndexBuffer: u32('sorted.bin'), shDegree: meta.shDegree,
}, null, null);
var target = new Float32Array(C.TILE_SPLAT_FRAME_UNIFORM_BYTES / 4);
C.writeTileSplatFrameUniforms(target, f32('viewproj.bin'), plan, 1.0, meta.shDegree, f32('campos.bin'), f32('view.bin'), f32('proj.bin'));
device.queue.writeBuffer(resources.frameUniformBuffer, 0, target);
var encoder = device.createCommandEncoder();
C.encodeFullComputeCompositorPipeline(encoder, resources, bindGroups);
device.queue.submit([bindGroups]);
fs.writeFileSync(path.join(dir, 'uniforms_js.bin'), Buffer.from(target.buffer));
fs.writeFileSync(path.join(dir, 'rgba_js.bin'), Buffer.from(resources.readRGBA().buffer));
fs.writeFileSync(path.join(dir, 'depth_js.bin'), Buffer.from(resources.readDepth().buffer));
fs.writeFileSync(path.join(dir, 'entries_js.bin'), Buffer.from(resources.debug('entries').buffer));
fs.writeFileSync(path.join(dir, 'keys_js.bin'), Buffer.from(resources.debug('keys20').buffer));
if (meta.gbuffer) {
  // the deferred consumers of the G-buffer, as the reference's encode* calls after the composite
  var px = meta.width * meta.height;
  fs.writeFileSync(path.join(dir, 'gbn_js.bin'), Buffer.from(resources.readOutput('gbufferNormal', new Uint32Array(px)).buffer));
  var opts = { recoveryMode: meta.recoveryMode };
  C.encodeScreenSpaceNormals(encoder, resources, opts);
  fs.writeFileSync(path.join(dir, 'ssn_js.bin'), Buffer.from(resources.readOutput('filteredNormal', new Uint32Array(px)).buffer));
  C.encodeBilateralNormalFilter(encoder, resources, opts);
  fs.writeFileSync(path.join(dir, 'bil_js.bin'), Buffer.from(resources.readOutput('gbufferNormal', new Uint32Array(px)).buffer));
}
console.log(JSON.stringify(resources.stats()));
resources.destroy();
