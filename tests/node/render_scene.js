'use strict';
// Renders one frame through the JS facade in the reference's call order and call shape
// (src/splatRenderer.ts:901-961 load, :1029-1040 frame) from binaries written by the Python test:
// the bind names exactly the reference's keys (no SH degree: the facade takes it from the shData
// length, the frame uniforms carry it), the outputs are textures from device.createTexture
// (rgba16float colour, r32float depth, and with the G-buffer r32uint normal / rgba32uint material)
// that queue.submit fills.  Writes the outputs and intermediates back next to the inputs.
var fs = require('fs');
var path = require('path');
var pkg = process.argv[2], dir = process.argv[3];
var C = require(path.join(pkg, 'node', 'gpuTileSplatCompositor.js'));
var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function f32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }
function save(name, a) { fs.writeFileSync(path.join(dir, name), Buffer.from(a.buffer, a.byteOffset, a.byteLength)); }
var device = C.createDevice();
var plan = C.planTileSplatCompositor({ viewportWidth: meta.width, viewportHeight: meta.height, tileSizePx: 16, splatCount: meta.count });
var resources = C.createTileSplatCompositor(device, plan, { debug: true, gbuffer: !!meta.gbuffer });
var size = [meta.width, meta.height];
var colour = device.createTexture({ label: 'compute_output', size: size, format: 'rgba16float' });
var gb = { depth: device.createTexture({ label: 'gbuffer_depth', size: size, format: 'r32float' }) };
if (meta.gbuffer) {
  gb.normal = device.createTexture({ label: 'gbuffer_normal', size: size, format: 'r32uint' });
  gb.material = device.createTexture({ label: 'gbuffer_material', size: size, format: 'rgba32uint' });
}
var bindGroups = C.createTileSplatBindGroups(device, resources, {
  positionBuffer: f32('positions.bin'), scaleBuffer: f32('scales.bin'), rotationBuffer: f32('rotations.bin'),
  opacityBuffer: f32('opacities.bin'), materialBuffer: u32('materials.bin'), normalBuffer: undefined,
  shDataBuffer: f32('shdata.bin'), sortedIndexBuffer: u32('sorted.bin'),
}, colour, gb);
var target = new Float32Array(C.TILE_SPLAT_FRAME_UNIFORM_BYTES / 4);
C.writeTileSplatFrameUniforms(target, f32('viewproj.bin'), plan, 1.0, meta.shDegree, f32('campos.bin'), f32('view.bin'), f32('proj.bin'));
device.queue.writeBuffer(resources.frameUniformBuffer, 0, target);
var encoder = device.createCommandEncoder();
C.encodeFullComputeCompositorPipeline(encoder, resources, bindGroups);
device.queue.submit([bindGroups]);
save('uniforms_js.bin', target);
save('rgba_js.bin', resources.readRGBA());
save('rgba16_tex.bin', colour.data);
save('depth_js.bin', resources.readDepth());
save('depth_tex.bin', gb.depth.data);
save('entries_js.bin', resources.debug('entries'));
save('keys_js.bin', resources.debug('keys20'));
if (meta.gbuffer) {
  save('gbn_tex.bin', gb.normal.data);
  save('gbm_tex.bin', gb.material.data);
  // the deferred consumers of the G-buffer, as the reference's encode* calls after the composite
  var px = meta.width * meta.height;
  save('gbn_js.bin', resources.readOutput('gbufferNormal', new Uint32Array(px)));
  var opts = { recoveryMode: meta.recoveryMode };
  C.encodeScreenSpaceNormals(encoder, resources, opts);
  save('ssn_js.bin', resources.readOutput('filteredNormal', new Uint32Array(px)));
  C.encodeBilateralNormalFilter(encoder, resources, opts);
  save('bil_js.bin', resources.readOutput('gbufferNormal', new Uint32Array(px)));
  // then the lighting pass, with encode's arguments by name (src/splatRenderer.ts:1115-1139)
  C.encodeDeferredLighting(encoder, resources, {
    lightDirection: [0.3, -0.8, -0.5], lightIntensity: 2.5, ambientIntensity: 0.12, exposure: 1.2,
    emissiveIntensity: 3.0, emissiveThreshold: 0.05, roughnessCurve: { contrast: 1.2, brightness: 0.0, gamma: 0.9 },
  });
  save('lit_js.bin', resources.readOutput('lit', new Uint16Array(px * 4)));
}
console.log(JSON.stringify(resources.stats()));
resources.destroy();
