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
// usage: node cpu_path.js <dir> [rowStep] [frames] [workers]
//   dir      the binaries bench.py writes (positions, scales, rotations, opacities, shdata,
//            uniforms, meta.json)
//   rowStep  composite tile rows ty % rowStep == 0 only (1 = the whole frame)
//   frames   timed frames after one untimed warm-up frame (default 1; 0 = no warm-up, one frame)
//   workers  1 = one thread (default); W > 1 = worker_threads: the projection, the per-tile sorts
//            and the compositing are shared out to W workers in dynamically claimed items (4096
//            splats, one tile, one tile row), the depth sort and the count / scan / scatter stay
//            on the main thread as in the single-thread path.
// Prints one JSON line: the median of each phase's seconds over the timed frames, each frame's
// total, the worker count and the host CPU model.

var fs = require('fs');
var path = require('path');
var os = require('os');
var wt = require('worker_threads');

var C1 = 0.4886025119029199, C2 = [1.0925484305920792, 1.0925484305920792, 0.31539156525252005, 1.0925484305920792,
  0.5462742152960396], C3 = [0.5900435899266435, 2.890611442640554, 0.4570457994644658, 0.3731763325901154,
  0.4570457994644658, 1.445305721320277, 0.5900435899266435];

function now() { return Number(process.hrtime.bigint()) / 1e9; }

// binary16 quantisation (round to nearest even), as pack2x16float does
var f32buf = new Float32Array(1), u32buf = new Uint32Array(f32buf.buffer);
function f16round(x) {
  f32buf[0] = x;
  var b = u32buf[0], s = b >>> 31, e = (b >>> 23) & 255;
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

var PROJ_BLOCK = 4096;

// Typed arrays on SharedArrayBuffers (shared with the workers) or plain buffers (one thread).
function alloc(ctor, n, shared) {
  n = Math.max(1, n);
  return shared ? new ctor(new SharedArrayBuffer(n * ctor.BYTES_PER_ELEMENT)) : new ctor(n);
}

// The per-frame arrays of a scene of N splats at W x H (reused across frames).
function frameArrays(N, W, H, shared) {
  var f = {};
  ['cx', 'cy', 'k0', 'k1', 'k2', 'op', 'cr', 'cg', 'cb', 'dep', 'ld'].forEach(function (k) {
    f[k] = alloc(Float64Array, N, shared);
  });
  f.sorted = alloc(Uint32Array, N, shared);
  f.box = alloc(Uint32Array, N, shared);
  f.rgba = alloc(Float32Array, W * H * 4, shared);
  f.counter = alloc(Int32Array, 1, shared);
  return f;
}

// ---- the three parallel phases: each processes work items [item] of its kind ----------------

// Projection cache of the sort ranks of block 'blk' (gpu_project_splats.wgsl:112-358).
function projectBlock(S, f, blk) {
  var N = S.N, W = S.W, H = S.H, U = S.U, vp = U.subarray(0, 16), view = U.subarray(32, 48);
  var fx = U[48], fy = U[49], cam = [U[28], U[29], U[30]], pos = S.pos, rot = S.rot, lscale = S.lscale;
  var shd = S.shd, opac = S.opac, deg = S.deg, stride = S.stride;
  var mtx = Math.max(S.tilesX, 1) - 1, mty = Math.max(S.tilesY, 1) - 1, ts = 16;
  var rEnd = Math.min(N, (blk + 1) * PROJ_BLOCK);
  for (var r = blk * PROJ_BLOCK; r < rEnd; r++) {
    var s = f.sorted[r], p = s * 3, X = pos[p], Y = pos[p + 1], Z = pos[p + 2];
    f.box[r] = 0xFFFFFFFF;
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
    f.k0[r] = cyy * inv; f.k1[r] = -cxy * inv; f.k2[r] = cxx * inv;
    f.op[r] = f16round(Math.min(Math.max(opac[s] * aa, 0), 0.99));
    f.cx[r] = px; f.cy[r] = py;
    var dep = (vp[2] * X + vp[6] * Y + vp[10] * Z + vp[14]) / sw;
    f.dep[r] = dep;
    f.ld[r] = Math.log(Math.max(dep, 1e-6));
    var rc = Math.min(rad, ts * 8);
    var x0 = Math.min(Math.max(Math.floor(Math.max(px - rc, 0) / ts), 0), mtx);
    var y0 = Math.min(Math.max(Math.floor(Math.max(py - rc, 0) / ts), 0), mty);
    var x1 = Math.min(Math.floor(Math.max((Math.min(px + rc, W) - 1e-9) / ts, 0)), mtx);
    var y1 = Math.min(Math.floor(Math.max((Math.min(py + rc, H) - 1e-9) / ts, 0)), mty);
    f.box[r] = (x0 | (y0 << 8) | (x1 << 16) | (y1 << 24)) >>> 0;
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
            C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
            C3[6] * x * (xx - 3 * yy)];
          for (c = 0; c < 3; c++) for (jj = 0; jj < 7; jj++) col[c] += shd[base + 27 + 3 * jj + c] * b3[jj];
        }
      }
    }
    f.cr[r] = f16round(Math.min(Math.max(col[0], 0), 1)); f.cg[r] = f16round(Math.min(Math.max(col[1], 0), 1));
    f.cb[r] = f16round(Math.min(Math.max(col[2], 0), 1));
  }
}

// One tile's sort: <= 4096 entries on (q20, local index); more: 128 log-depth buckets (stable),
// then consecutive 4096-entry chunks, each sorted the same way.
function sortTile(f, L, m, scratch) {
  var entries = L.entries, c0 = L.counts[m], st = L.offs[m], ld = f.ld, dep = f.dep;
  var keybuf = scratch.keybuf, tmp = scratch.tmp;
  function sortSeg(start, n, lmin, scale) {
    for (var i = 0; i < n; i++) {
      var q = Math.min(Math.max(Math.floor((ld[entries[start + i]] - lmin) * scale * 8192), 0), 1048575);
      keybuf[i] = ((q << 12) | i) >>> 0;
    }
    var k = keybuf.subarray(0, n); k.sort();
    for (i = 0; i < n; i++) tmp[i] = entries[start + (k[i] & 4095)];
    for (i = 0; i < n; i++) entries[start + i] = tmp[i];
  }
  var dmin = Infinity, dmax = -Infinity, i;
  for (i = 0; i < c0; i++) { var dd = dep[entries[st + i]]; if (dd < dmin) dmin = dd; if (dd > dmax) dmax = dd; }
  var lmin = Math.log(Math.max(dmin, 1e-6)), lr = Math.log(Math.max(dmax, 1e-6)) - lmin, scale = lr < 1e-10 ? 0 : 128 / lr;
  if (c0 <= 4096) { sortSeg(st, c0, lmin, scale); return; }
  var bc = new Uint32Array(129), bk = new Uint8Array(c0), ordered = new Uint32Array(c0), b;
  for (i = 0; i < c0; i++) { b = Math.min(Math.max(Math.floor((ld[entries[st + i]] - lmin) * scale), 0), 127); bk[i] = b; bc[b + 1]++; }
  for (b = 0; b < 128; b++) bc[b + 1] += bc[b];
  for (i = 0; i < c0; i++) ordered[bc[bk[i]]++] = entries[st + i];
  entries.set(ordered, st);
  for (var cs = 0; cs < c0; cs += 4096) sortSeg(st + cs, Math.min(4096, c0 - cs), lmin, scale);
}

// Front-to-back compositing of tile row ty (gpu_tile_splat_composite.wgsl:162-370).
function compositeRow(S, f, L, ty) {
  var W = S.W, H = S.H, entries = L.entries, rgba = f.rgba;
  for (var tx = 0; tx < S.tilesX; tx++) {
    var mm = morton(tx, ty), s0 = L.offs[mm], n0 = L.counts[mm];
    for (var py = ty * 16; py < Math.min(ty * 16 + 16, H); py++) {
      for (var px = tx * 16; px < Math.min(tx * 16 + 16, W); px++) {
        var T = 1, R0 = 0, G0 = 0, B0 = 0, qx = px + 0.5, qy = py + 0.5;
        for (var i = 0; i < n0 && T >= 0.001; i++) {
          var e = entries[s0 + i], ddx = qx - f.cx[e], ddy = qy - f.cy[e];
          var pw = -0.5 * f.k0[e] * ddx * ddx - f.k1[e] * ddx * ddy - 0.5 * f.k2[e] * ddy * ddy;
          if (pw <= -4) continue;
          var al = Math.min(0.99, f.op[e] * Math.exp(pw));
          if (al <= 1 / 255) continue;
          var w = al * T;
          R0 += f.cr[e] * w; G0 += f.cg[e] * w; B0 += f.cb[e] * w; T *= 1 - al;
        }
        var o4 = (py * W + px) * 4;
        rgba[o4] = R0 + T * 0.02; rgba[o4 + 1] = G0 + T * 0.02; rgba[o4 + 2] = B0 + T * 0.04; rgba[o4 + 3] = 1 - T;
      }
    }
  }
}

function newScratch() { return { keybuf: new Uint32Array(4096), tmp: new Uint32Array(4096) }; }

// Claims items from the shared counter until none is left.
function drain(counter, n, fn) {
  for (;;) {
    var i = Atomics.add(counter, 0, 1);
    if (i >= n) return;
    fn(i);
  }
}

// ---- worker side -------------------------------------------------------------------------------
if (!wt.isMainThread) {
  var S = wt.workerData.scene, f = wt.workerData.frame, L = null, scratch = newScratch();
  wt.parentPort.on('message', function (msg) {
    if (msg.lists) L = msg.lists;
    if (msg.phase === 'project') drain(f.counter, msg.n, function (blk) { projectBlock(S, f, blk); });
    else if (msg.phase === 'sort') drain(f.counter, msg.n, function (i) { sortTile(f, L, msg.tiles[i], scratch); });
    else if (msg.phase === 'composite') drain(f.counter, msg.n, function (i) { compositeRow(S, f, L, i * msg.rowStep); });
    wt.parentPort.postMessage('done');
  });
  return;
}

// ---- main thread -------------------------------------------------------------------------------
function loadScene(dir, shared) {
  var meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  function load(name, ctor) {
    var b = fs.readFileSync(path.join(dir, name));
    var a = alloc(ctor, b.length / ctor.BYTES_PER_ELEMENT, shared);
    new Uint8Array(a.buffer).set(new Uint8Array(b.buffer, b.byteOffset, b.length));
    return a;
  }
  var U = load('uniforms.bin', Float32Array), Uu = new Uint32Array(U.buffer, U.byteOffset, 64);
  var K = (meta.shDegree + 1) * (meta.shDegree + 1) - 1;
  return {
    meta: meta, N: meta.count, W: meta.width, H: meta.height, deg: meta.shDegree, stride: 3 + 3 * K + 3,
    pos: load('positions.bin', Float32Array), lscale: load('scales.bin', Float32Array),
    rot: load('rotations.bin', Float32Array), opac: load('opacities.bin', Float32Array),
    shd: load('shdata.bin', Float32Array), U: U, tilesX: Uu[20], tilesY: Uu[21],
  };
}

function makePool(S, f, workers) {
  if (workers <= 1) return null;
  var pool = [];
  for (var i = 0; i < workers; i++) {
    pool.push(new wt.Worker(__filename, { workerData: { scene: S, frame: f } }));
  }
  return pool;
}

// Runs n items of a phase: inline with no pool, else on every worker (items claimed dynamically).
function runPhase(pool, f, msg, inline) {
  f.counter[0] = 0;
  if (!pool) {
    for (var i = 0; i < msg.n; i++) inline(i);
    return Promise.resolve();
  }
  return Promise.all(pool.map(function (w) {
    return new Promise(function (resolve) {
      w.once('message', resolve);
      w.postMessage(msg);
    });
  }));
}

async function frame(S, f, pool, rowStep, scratch) {
  var t = {}, t0 = now(), N = S.N, i, r;
  // 1. cpuDepthSort (one thread, as the reference)
  var view = S.U.subarray(32, 48), pos = S.pos;
  var keys = new Array(N);
  for (var id = 0; id < N; id++) {
    var o = id * 3;
    keys[id] = { id: id, depth: view[2] * pos[o] + view[6] * pos[o + 1] + view[10] * pos[o + 2] + view[14] };
  }
  keys.sort(function (a, b) { return a.depth - b.depth || a.id - b.id; });
  for (i = 0; i < N; i++) f.sorted[i] = keys[i].id;
  keys = null;
  t.cpuDepthSort_s = now() - t0;

  // 2a. projection cache
  t0 = now();
  var nblk = Math.ceil(N / PROJ_BLOCK);
  await runPhase(pool, f, { phase: 'project', n: nblk }, function (blk) { projectBlock(S, f, blk); });
  t.project_s = now() - t0;

  // 2b. count, scan, canonical scatter (main thread), per-tile sort
  t0 = now();
  var M = morton(S.tilesX - 1, S.tilesY - 1) + 1, shared = !!pool;
  var counts = alloc(Uint32Array, M, shared), offs = alloc(Uint32Array, M, shared), cursor = new Uint32Array(M);
  var large = [], box = f.box, bd, tx, ty;
  for (r = 0; r < N; r++) {
    bd = box[r];
    if (bd === 0xFFFFFFFF) continue;
    var bx0 = bd & 255, by0 = (bd >>> 8) & 255, bx1 = (bd >>> 16) & 255, by1 = bd >>> 24;
    if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > 16) large.push(r);
    for (ty = by0; ty <= by1; ty++) for (tx = bx0; tx <= bx1; tx++) counts[morton(tx, ty)]++;
  }
  var R_ = 0, m;
  for (m = 0; m < M; m++) { offs[m] = R_; R_ += counts[m]; }
  var entries = alloc(Uint32Array, R_, shared);
  var scatter = function (r) {
    var bd = box[r], bx0 = bd & 255, by0 = (bd >>> 8) & 255, bx1 = (bd >>> 16) & 255, by1 = bd >>> 24;
    for (var ty = by0; ty <= by1; ty++) for (var tx = bx0; tx <= bx1; tx++) { var t = morton(tx, ty); entries[offs[t] + cursor[t]++] = r; }
  };
  for (r = 0; r < N; r++) {
    bd = box[r];
    if (bd === 0xFFFFFFFF) continue;
    if ((((bd >>> 16) & 255) - (bd & 255) + 1) * ((bd >>> 24) - ((bd >>> 8) & 255) + 1) <= 16) scatter(r);
  }
  for (i = 0; i < large.length; i++) scatter(large[i]);
  var L = { entries: entries, counts: counts, offs: offs };
  // tiles in descending count order, so the long ones are claimed first
  var tl = [];
  for (m = 0; m < M; m++) if (counts[m]) tl.push(m);
  tl.sort(function (a, b) { return counts[b] - counts[a]; });
  var tiles = alloc(Uint32Array, tl.length, shared);
  tiles.set(tl);
  await runPhase(pool, f, { phase: 'sort', n: tiles.length, tiles: tiles, lists: L },
                 function (k) { sortTile(f, L, tiles[k], scratch); });
  t.tile_refs_s = now() - t0;

  // 3. JS compositing (tile rows ty % rowStep == 0)
  t0 = now();
  var rows = Math.ceil(S.tilesY / rowStep);
  await runPhase(pool, f, { phase: 'composite', n: rows, rowStep: rowStep, lists: L },
                 function (k) { compositeRow(S, f, L, k * rowStep); });
  t.composite_sample_s = now() - t0;
  t.composite_rows = rows;
  t.refs = R_;
  t.frame_s_est = t.cpuDepthSort_s + t.project_s + t.tile_refs_s + t.composite_sample_s * S.tilesY / rows;
  return t;
}

function median(a) {
  var s = a.slice().sort(function (x, y) { return x - y; });
  var n = s.length;
  return n % 2 ? s[(n - 1) / 2] : 0.5 * (s[n / 2 - 1] + s[n / 2]);
}

async function main() {
  var dir = process.argv[2];
  var rowStep = parseInt(process.argv[3] || '1', 10);
  var frames = parseInt(process.argv[4] || '1', 10);
  var workers = parseInt(process.argv[5] || '1', 10);
  var shared = workers > 1;
  var S = loadScene(dir, shared);
  var f = frameArrays(S.N, S.W, S.H, shared);
  var pool = makePool(S, f, workers);
  var scratch = newScratch();
  var runs = [];
  if (frames > 0) await frame(S, f, pool, rowStep, scratch);  // warm-up
  for (var k = 0; k < Math.max(1, frames); k++) runs.push(await frame(S, f, pool, rowStep, scratch));
  if (pool) pool.forEach(function (w) { w.terminate(); });
  var out = {};
  ['cpuDepthSort_s', 'project_s', 'tile_refs_s', 'composite_sample_s', 'frame_s_est'].forEach(function (key) {
    out[key] = median(runs.map(function (t) { return t[key]; }));
  });
  out.frames_s = runs.map(function (t) { return t.frame_s_est; });
  out.frames = runs.length;
  out.warmup = frames > 0 ? 1 : 0;
  out.composite_rows = runs[0].composite_rows;
  out.tile_rows = S.tilesY;
  out.refs = runs[0].refs;
  out.workers = workers;
  out.cpu_model = os.cpus()[0].model;
  out.cpus = os.cpus().length;
  var sum = 0;
  for (var i = 0; i < f.rgba.length; i += 4) sum += f.rgba[i] + f.rgba[i + 1] + f.rgba[i + 2];
  out.sampled_rgb_sum = sum;
  if (S.meta.writeRgba) fs.writeFileSync(path.join(dir, 'rgba_cpu.bin'), Buffer.from(f.rgba.buffer, 0, S.W * S.H * 16));
  console.log(JSON.stringify(out));
}

main().catch(function (e) { console.error(e && e.stack ? e.stack : String(e)); process.exitCode = 1; });
