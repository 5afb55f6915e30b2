'use strict';
// CPU-only checks of the JS facade (no GPU work): plan arithmetic of
// src/gpuTileSplatCompositor.ts:109-139, the 256-byte uniform layout, and that unsupported
// options raise.  Prints one JSON object; the Python test asserts on it.
var path = require('path');
var C = require(path.join(process.argv[2], 'node', 'gpuTileSplatCompositor.js'));
var out = {};
out.plans = [[640, 480, 50000], [1920, 1080, 1000000], [3840, 2160, 6000000]].map(function (c) {
  return C.planTileSplatCompositor({ viewportWidth: c[0], viewportHeight: c[1], tileSizePx: 16, splatCount: c[2] });
});
var plan = out.plans[0];
var t = new Float32Array(64);
var vp = new Float32Array(16); for (var i = 0; i < 16; i++) vp[i] = i + 0.5;
var view = new Float32Array(16); for (i = 0; i < 16; i++) view[i] = -i;
var proj = new Float32Array(16); proj[0] = 1.640625; proj[5] = 2.1875;
C.writeTileSplatFrameUniforms(t, vp, plan, 1.0, 3, [1, 2, 3], view, proj);
out.uniforms = Array.prototype.slice.call(new Uint32Array(t.buffer));
out.uniformBytes = C.TILE_SPLAT_FRAME_UNIFORM_BYTES;
var bad = Object.assign({}, plan, { tileSizePx: 8 });  // rejected before any GPU work
try { C.createTileSplatCompositor(null, bad, {}); out.badTile = 'accepted'; } catch (e) { out.badTile = String(e.message); }
out.version = C.version();
console.log(JSON.stringify(out));
