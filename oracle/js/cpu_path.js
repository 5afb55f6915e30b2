'use strict';
// cpu_path.js - the reference's CPU path for the compositor, restated in plain JavaScript
// (Node >= 12), for the CPU baseline timing.  TEST / BENCHMARK INFRASTRUCTURE ONLY.
//
//   1. cpuDepthSort: sortSplatIdsBackToFront (src/splatSort.ts:32-55): {id, depth} objects sorted
//      by view-space z, ties by id.
//   2. CPU tile-ref population: projection cache (src/shaders/gpu_project_splats.wgsl:112-358),
//      per-tile counts, exclusive scan, canonical scatter (small splats then large splats,
//      gpu_tile_splat_composite.wgsl:64-119, gpu_large_splat_scatter.wgsl:52-89), per-tile sort
//      on (20-bit log-depth key, local index) with the 128-bucket cascade above 4096 entries
//      (gpu_tile_depth_sort.wgsl, gpu_tile_bucket_sort.wgsl, gpu_tile_chunk_sort.wgsl).
//   3. JS compositing: front-to-back per pixel (gpu_tile_splat_composite.wgsl:162-370).
// Arithmetic is JavaScript double precision (colour and opacity quantised to binary16 as in the
// cache), so images agree with the fp32 path to rounding, not bit for bit.
//
// usage: node cpu_path.js <dir> <rowStep>   (dir holds the binaries bench.py writes)
// Prints one JSON line with per-phase seconds; the composite covers tile rows ty % rowStep == 0.

var fs = require('fs');
var path = require('path');

var dir = process.argv[2];
var rowStep = parseInt(process.argv[3] || '1', 10);
var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function f32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(name) { var b = fs.readFileSync(path.join(dir, name)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }

var pos = f32('positions.bin'), lscale = f32('scales.bin'), rot = f32('rotations.bin'), opac = f32('opacities.bin');
var shd = f32('shdata.bin'), U = f32('uniforms.bin'), Uu = new Uint32Array(U.buffer, U.byteOffset, 64);
var N = meta.count, W = meta.width, H = meta.height, deg = meta.shDegree;
var K = (deg + 1) * (deg + 1) - 1, stride = 3 + 3 * K + 3;
var tilesX = Uu[20], tilesY = Uu[21];

function now() { return Number(process.hrtime.bigint()) / 1e9; }

// binary16 quantisation (round to nearest even), as pack2x16float does
var f32buf = new Float32Array(1), u32buf = new Uint32Array(f32buf.buffer);
function f16round(x) {
  f32buf[0] = x;
  var b = u32buf[0], s = b >>> 31, e = (b >>> 23) & 255, m = b & 0x7FFFFF;
  if (e === 255) return x;
  var v = Math.abs(f32buf[0]);
  if (v >= 65520) return s ? -Infinity : Infinity;
  var ex = Math.floor(Math.log2(v));
  if (ex < -14) ex = -14;
  var ulp = Math.pow(2, ex - 10), q = v / ulp, r = Math.round(q);
  if (Math.abs(q - Math.floor(q) - 0.5) < 1e-12) r = (Math.floor(q) % 2 === 0) ? Math.floor(q) : Math.floor(q) + 1;
  var out = r * ulp;
  return s ? -out : out;
}

function morton(x, y) {
  var mx = x & 0xFFFF; mx = (mx | (mx << 8)) & 0x00FF00FF; mx = (mx | (mx << 4)) & 0x0F0F0F0F;
  mx = (mx | (mx << 2)) & 0x33333333; mx = (mx | (mx << 1)) & 0x55555555;
  var my = y & 0xFFFF; my = (my | (my << 8)) & 0x00FF00FF; my = (my | (my << 4)) & 0x0F0F0F0F;
  my = (my | (my << 2)) & 0x33333333; my = (my | (my << 1)) & 0x55555555;
  return (mx | (my << 1)) >>> 0;
}

var times = {};
var t0 = now();

// 1. cpuDepthSort
var view = U.subarray(32, 48);
var keys = new Array(N);
for (var id = 0; id < N; id++) {
  var o = id * 3;
  keys[id] = { id: id, depth: view[2] * pos[o] + view[6] * pos[o + 1] + view[10] * pos[o + 2] + view[14] };
}
keys.sort(function (a, b) { return a.depth - b.depth || a.id - b.id; });
var sorted = new Uint32Array(N);
for (var i = 0; i < N; i++) sorted[i] = keys[i].id;
keys = null;
times.cpuDepthSort_s = now() - t0;

// 2a. projection cache
t0 = now();
var vp = U.subarray(0, 16), fx = U[48], fy = U[49], cam = [U[28], U[29], U[30]];
var cx = new Float64Array(N), cy = new Float64Array(N), k0 = new Float64Array(N), k1 = new Float64Array(N),
  k2 = new Float64Array(N), op = new Float64Array(N), cr = new Float64Array(N), cg = new Float64Array(N),
  cb = new Float64Array(N), dep = new Float64Array(N), box = new Uint32Array(N);
var C1 = 0.4886025119029199, C2 = [1.0925484305920792, 1.0925484305920792, 0.31539156525252005, 1.0925484305920792,
  0.5462742152960396], C3 = [0.5900435899266435, 2.890611442640554, 0.4570457994644658, 0.3731763325901154,
  0.4570457994644658, 1.445305721320277, 0.5900435899266435];
var mtx = Math.max(tilesX, 1) - 1, mty = Math.max(tilesY, 1) - 1, ts = 16;
for (var r = 0; r < N; r++) {
  var s = sorted[r], p = s * 3, X = pos[p], Y = pos[p + 1], Z = pos[p + 2];
  box[r] = 0xFFFFFFFF;
  var cw = vp[3] * X + vp[7] * Y + vp[11] * Z + vp[15];
  if (cw <= 0) continue;
  var sw = Math.max(cw, 1e-4);
  var nx = (vp[0] * X + vp[4] * Y + vp[8] * Z + vp[12]) / sw, ny = (vp[1] * X + vp[5] * Y + vp[9] * Z + vp[13]) / sw;
  var px = (nx * 0.5 + 0.5) * W, py = (1 - (ny * 0.5 + 0.5)) * H;
  var q = s * 4, qw = rot[q], qx = rot[q + 1], qy = rot[q + 2], qz = rot[q + 3];
  var ql = Math.max(Math.sqrt(qw * qw + qx * qx + qy * qy + qz * qz), 1e-6);
  qw /= ql; qx /= ql; qy /= ql; qz /= ql;
  // rotation matrix columns = rotated unit axes, scaled
  var R = [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy + qw * qz), 2 * (qx * qz - qw * qy),
    2 * (qx * qy - qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz + qw * qx),
    2 * (qx * qz + qw * qy), 2 * (qy * qz - qw * qx), 1 - 2 * (qx * qx + qy * qy)];
  var sc = [Math.exp(lscale[p]), Math.exp(lscale[p + 1]), Math.exp(lscale[p + 2])];
  var vx = view[0] * X + view[4] * Y + view[8] * Z + view[12], vy = view[1] * X + view[5] * Y + view[9] * Z + view[13];
  var vz = Math.min(view[2] * X + view[6] * Y + view[10] * Z + view[14], -0.001);
  var j1x = fx / vz, j1y = fy / vz, j2x = (-j1x / vz) * vx, j2y = (-j1y / vz) * vy;
  var t0x = j1x * view[0] + j2x * view[2], t0y = j1x * view[4] + j2x * view[6], t0z = j1x * view[8] + j2x * view[10];
  var t1x = j1y * view[1] + j2y * view[2], t1y = j1y * view[5] + j2y * view[6], t1z = j1y * view[9] + j2y * view[10];
  var cxx = 0, cxy = 0, cyy = 0;
  for (var a = 0; a < 3; a++) {
    var ax = R[a * 3] * sc[a], ay = R[a * 3 + 1] * sc[a], az = R[a * 3 + 2] * sc[a];
    var b0 = ax * t0x + ay * t0y + az * t0z, b1 = ax * t1x + ay * t1y + az * t1z;
    cxx += b0 * b0; cxy += b0 * b1; cyy += b1 * b1;
  }
  var detO = cxx * cyy - cxy * cxy; cxx += 0.3; cyy += 0.3;
  var detB = cxx * cyy - cxy * cxy, aa = Math.sqrt(Math.max(detO / Math.max(detB, 1e-6), 0));
  var mid = 0.5 * (cxx + cyy), lam = mid + Math.sqrt(Math.max(mid * mid - Math.max(detB, 0.1), 0.1));
  var rad = Math.ceil(3 * Math.sqrt(lam));
  if (!(rad > 0)) continue;
  var inv = 1 / Math.max(detB, 1e-6);
  k0[r] = cyy * inv; k1[r] = -cxy * inv; k2[r] = cxx * inv;
  op[r] = f16round(Math.min(Math.max(opac[s] * aa, 0), 0.99));
  cx[r] = px; cy[r] = py; dep[r] = (vp[2] * X + vp[6] * Y + vp[10] * Z + vp[14]) / sw;
  var rc = Math.min(rad, ts * 8);
  var x0 = Math.min(Math.max(Math.floor(Math.max(px - rc, 0) / ts), 0), mtx), y0 = Math.min(Math.max(Math.floor(Math.max(py - rc, 0) / ts), 0), mty);
  var x1 = Math.min(Math.floor(Math.max((Math.min(px + rc, W) - 1e-9) / ts, 0)), mtx), y1 = Math.min(Math.floor(Math.max((Math.min(py + rc, H) - 1e-9) / ts, 0)), mty);
  box[r] = (x0 | (y0 << 8) | (x1 << 16) | (y1 << 24)) >>> 0;
  // SH colour, WGSL sign convention
  var dx = cam[0] - X, dy = cam[1] - Y, dz = cam[2] - Z, dl = Math.sqrt(dx * dx + dy * dy + dz * dz);
  var x = dx / dl, y = dy / dl, z = dz / dl, base = s * stride, col = [shd[base], shd[base + 1], shd[base + 2]];
  if (deg >= 1) {
    for (var c = 0; c < 3; c++) col[c] += C1 * (shd[base + 3 + c] * y + shd[base + 6 + c] * z + shd[base + 9 + c] * x);
    if (deg >= 2) {
      var xx = x * x, yy = y * y, zz = z * z;
      var b2 = [C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z, C2[4] * (xx - yy)];
      for (c = 0; c < 3; c++) for (var jj = 0; jj < 5; jj++) col[c] += shd[base + 12 + 3 * jj + c] * b2[jj];
      if (deg >= 3) {
        var b3 = [C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
          C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy), C3[6] * x * (xx - 3 * yy)];
        for (c = 0; c < 3; c++) for (jj = 0; jj < 7; jj++) col[c] += shd[base + 27 + 3 * jj + c] * b3[jj];
      }
    }
  }
  cr[r] = f16round(Math.min(Math.max(col[0], 0), 1)); cg[r] = f16round(Math.min(Math.max(col[1], 0), 1));
  cb[r] = f16round(Math.min(Math.max(col[2], 0), 1));
}
times.project_s = now() - t0;

// 2b. count, scan, canonical scatter, per-tile sort
t0 = now();
var M = morton(tilesX - 1, tilesY - 1) + 1;
var counts = new Uint32Array(M), offs = new Uint32Array(M), cursor = new Uint32Array(M);
var large = [];
for (r = 0; r < N; r++) {
  var bd = box[r];
  if (bd === 0xFFFFFFFF) continue;
  var bx0 = bd & 255, by0 = (bd >>> 8) & 255, bx1 = (bd >>> 16) & 255, by1 = bd >>> 24;
  if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > 16) large.push(r);
  for (var ty = by0; ty <= by1; ty++) for (var tx = bx0; tx <= bx1; tx++) counts[morton(tx, ty)]++;
}
var R_ = 0;
for (var m = 0; m < M; m++) { offs[m] = R_; R_ += counts[m]; }
var entries = new Uint32Array(R_);
function scatter(r) {
  var bd = box[r], bx0 = bd & 255, by0 = (bd >>> 8) & 255, bx1 = (bd >>> 16) & 255, by1 = bd >>> 24;
  for (var ty = by0; ty <= by1; ty++) for (var tx = bx0; tx <= bx1; tx++) { var t = morton(tx, ty); entries[offs[t] + cursor[t]++] = r; }
}
for (r = 0; r < N; r++) {
  bd = box[r];
  if (bd === 0xFFFFFFFF) continue;
  if ((((bd >>> 16) & 255) - (bd & 255) + 1) * ((bd >>> 24) - ((bd >>> 8) & 255) + 1) <= 16) scatter(r);
}
for (i = 0; i < large.length; i++) scatter(large[i]);
var ld = new Float64Array(N);
for (r = 0; r < N; r++) ld[r] = Math.log(Math.max(dep[r], 1e-6));
var keybuf = new Uint32Array(4096), tmp = new Uint32Array(4096);
function sortSeg(start, n, lmin, scale) {
  for (var i = 0; i < n; i++) {
    var q = Math.min(Math.max(Math.floor((ld[entries[start + i]] - lmin) * scale * 8192), 0), 1048575);
    keybuf[i] = ((q << 12) | i) >>> 0;
  }
  var k = keybuf.subarray(0, n); k.sort();
  for (i = 0; i < n; i++) tmp[i] = entries[start + (k[i] & 4095)];
  for (i = 0; i < n; i++) entries[start + i] = tmp[i];
}
for (m = 0; m < M; m++) {
  var c0 = counts[m], st = offs[m];
  if (!c0) continue;
  var dmin = Infinity, dmax = -Infinity;
  for (i = 0; i < c0; i++) { var dd = dep[entries[st + i]]; if (dd < dmin) dmin = dd; if (dd > dmax) dmax = dd; }
  var lmin = Math.log(Math.max(dmin, 1e-6)), lr = Math.log(Math.max(dmax, 1e-6)) - lmin, scale = lr < 1e-10 ? 0 : 128 / lr;
  if (c0 <= 4096) { sortSeg(st, c0, lmin, scale); continue; }
  var bc = new Uint32Array(129), bk = new Uint8Array(c0), ordered = new Uint32Array(c0);
  for (i = 0; i < c0; i++) { var b = Math.min(Math.max(Math.floor((ld[entries[st + i]] - lmin) * scale), 0), 127); bk[i] = b; bc[b + 1]++; }
  for (b = 0; b < 128; b++) bc[b + 1] += bc[b];
  for (i = 0; i < c0; i++) ordered[bc[bk[i]]++] = entries[st + i];
  entries.set(ordered, st);
  for (var cs = 0; cs < c0; cs += 4096) sortSeg(st + cs, Math.min(4096, c0 - cs), lmin, scale);
}
times.tile_refs_s = now() - t0;

// 3. JS compositing (sampled tile rows)
t0 = now();
var rgba = new Float32Array(W * H * 4), rows = 0, sum = 0;
for (var tyy = 0; tyy < tilesY; tyy += rowStep) {
  rows++;
  for (var txx = 0; txx < tilesX; txx++) {
    var mm = morton(txx, tyy), s0 = offs[mm], n0 = counts[mm];
    for (var py_ = tyy * 16; py_ < Math.min(tyy * 16 + 16, H); py_++) {
      for (var px_ = txx * 16; px_ < Math.min(txx * 16 + 16, W); px_++) {
        var T = 1, R0 = 0, G0 = 0, B0 = 0, qx_ = px_ + 0.5, qy_ = py_ + 0.5;
        for (i = 0; i < n0 && T >= 0.001; i++) {
          var e = entries[s0 + i], ddx = qx_ - cx[e], ddy = qy_ - cy[e];
          var pw = -0.5 * k0[e] * ddx * ddx - k1[e] * ddx * ddy - 0.5 * k2[e] * ddy * ddy;
          if (pw <= -4) continue;
          var al = Math.min(0.99, op[e] * Math.exp(pw));
          if (al <= 1 / 255) continue;
          var w = al * T;
          R0 += cr[e] * w; G0 += cg[e] * w; B0 += cb[e] * w; T *= 1 - al;
        }
        var o4 = (py_ * W + px_) * 4;
        rgba[o4] = R0 + T * 0.02; rgba[o4 + 1] = G0 + T * 0.02; rgba[o4 + 2] = B0 + T * 0.04; rgba[o4 + 3] = 1 - T;
        sum += rgba[o4] + rgba[o4 + 1] + rgba[o4 + 2];
      }
    }
  }
}
times.composite_sample_s = now() - t0;
times.composite_rows = rows;
times.tile_rows = tilesY;
times.refs = R_;
times.frame_s_est = times.cpuDepthSort_s + times.project_s + times.tile_refs_s + times.composite_sample_s * tilesY / rows;
times.sampled_rgb_sum = sum;
if (meta.writeRgba) fs.writeFileSync(path.join(dir, 'rgba_cpu.bin'), Buffer.from(rgba.buffer));
console.log(JSON.stringify(times));
