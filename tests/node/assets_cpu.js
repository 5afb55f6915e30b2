'use strict';
// Decodes the assets named in <dir>/assets.json with the facade's loaders (no GPU work) and
// writes every decoded array and the packed bind buffers next to them, for the Python test.
var fs = require('fs');
var path = require('path');
var pkg = process.argv[2], dir = process.argv[3];
var A = require(path.join(pkg, 'node', 'splatAssets.js'));
var jobs = JSON.parse(fs.readFileSync(path.join(dir, 'assets.json'), 'utf8'));
var summary = {};
jobs.forEach(function (job) {
  var attrs;
  try {
    attrs = job.kind === 'ply' ? A.decodeLocalPlySplatPayload(path.basename(job.file), fs.readFileSync(job.file))
                               : A.loadFirstSmokeSplatPayload(job.file);
  } catch (e) {
    summary[job.name] = { error: String(e.message) };
    return;
  }
  var up = A.uploadSplatAttributeBuffers(null, attrs);
  var arrays = { positions: attrs.positions, colors: attrs.colors, opacities: attrs.opacities, radii: attrs.radii,
    scales: attrs.scales, rotations: attrs.rotations, originalIds: attrs.originalIds, materials: up.materialBuffer,
    shData: up.shDataBuffer };
  if (attrs.sh) arrays.sh = attrs.sh.coefficients;
  if (attrs.normals) arrays.normals = attrs.normals;
  Object.keys(arrays).forEach(function (k) {
    var a = arrays[k];
    fs.writeFileSync(path.join(dir, job.name + '.' + k + '.bin'), Buffer.from(a.buffer, a.byteOffset, a.byteLength));
  });
  summary[job.name] = { count: attrs.count, sourceKind: attrs.sourceKind, bounds: attrs.bounds,
    shDegree: up.shDegree, arrays: Object.keys(arrays), framing: A.framingFromBounds(attrs.bounds) };
});
console.log(JSON.stringify(summary));
