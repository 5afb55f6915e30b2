// k_project.hip - projection cache + tile count + large-splat classification, one pass.
//
// Reference: project_splats (src/shaders/gpu_project_splats.wgsl:112-358),
// classify_large_splats (gpu_classify_large_splats.wgsl:33-56) and the per-splat footprint of
// count_tile_refs (gpu_tile_splat_composite.wgsl:64-84).
//
// One thread per sort rank; 256 threads per workgroup; a workgroup never straddles two frames.
// The splat is gathered through sortedIndices (the only uncoalesced reads); the projected record
// is written as two coalesced 32-byte records plus the tile box and the depth bits.  Instead of
// atomically counting tiles, each workgroup reduces (small-splat pairs, large-splat pairs, large
// splats, visible splats) into one uint4, and the (small, large) counts of 8 x 8-tile cells the
// boxes touch into another: the scans of those block sums give every splat its offsets in
// canonical order with no atomics.
//
// Arithmetic is IEEE fp32 with no contraction, in the operation order of the WGSL, so the
// record is bit-identical to the CPU restatement.
#include "hdsr_internal.h"
#include "proj_common.h"

namespace hdsr {
namespace {

__device__ __forceinline__ void oct_encode(F3 n, float* ox, float* oy) {
  const float sum = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
  float px = n.x / sum, py = n.y / sum;
  if (n.z < 0.0f) {
    const float sx = px >= 0.0f ? 1.0f : -1.0f;
    const float sy = py >= 0.0f ? 1.0f : -1.0f;
    const float nx = (1.0f - fabsf(py)) * sx;
    const float ny = (1.0f - fabsf(px)) * sy;
    px = nx;
    py = ny;
  }
  *ox = px;
  *oy = py;
}

// kPacked: the scene is read from packed records (one or two 128-byte lines per splat) instead
// of the caller's attribute arrays.
template <bool kPacked>
__global__ __launch_bounds__(256) void k_project(Bufs b) {
  const uint32_t blk = blockIdx.x;
  const uint32_t f = frame_of(b.block_frame_start, b.num_frames, blk);
  const FrameDesc& F = b.frames[f];
  const uint32_t g = blk * kSplatBlock + threadIdx.x;
  const uint32_t r = g - F.splat_base;

  uint32_t n_small = 0, n_large = 0, is_large = 0, vis = 0, c_small = 0, c_large = 0;
  uint32_t bounds = HDSR_SENTINEL, depth_bits = HDSR_SENTINEL;
  uint4 h0 = make_uint4(0, 0, 0, 0), h1 = make_uint4(0, 0, 0, 0);
  uint4 c0 = make_uint4(0, 0, 0, 0), c1 = make_uint4(0, 0, 0, 0);

  const uint32_t id = (r < F.n) ? F.sorted[r] : 0xFFFFFFFFu;
  if (id < F.n) {
    const float4* R = kPacked ? F.packed + (size_t)id * F.pk_stride4 : nullptr;
    // packed: the whole record (at most 16 float4) is loaded up front, so both of its 128-byte
    // lines are requested together and consumed while resident (loading the SH part where it is
    // used lets the lines be evicted and fetched a second time); every index below is a constant
    float4 rv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) rv[j] = make_float4(0, 0, 0, 0);
    if (kPacked) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < 8 || F.pk_stride4 > 8u) rv[j] = R[j];
    }
    const float4 r0 = rv[0], r1 = rv[1], r2 = rv[2];
    const F3 center = kPacked ? f3(r1.x, r1.y, r1.z) : f3(F.pos[id * 3], F.pos[id * 3 + 1], F.pos[id * 3 + 2]);
    float clip[4];
    mat4_point(F.vp, center, clip);
    if (clip[3] > 0.0f) {
      const float safe_w = fmaxf(clip[3], HDSR_MIN_CLIP_W);
      const float ndc_x = clip[0] / safe_w, ndc_y = clip[1] / safe_w;
      const float px = (ndc_x * 0.5f + 0.5f) * F.vw;
      const float py = (1.0f - (ndc_y * 0.5f + 0.5f)) * F.vh;
      const float depth = clip[2] / safe_w;

      const float4 q4 = kPacked ? r0 : *reinterpret_cast<const float4*>(F.rot + (size_t)id * 4);
      const float qlen = sqrtf(((q4.x * q4.x + q4.y * q4.y) + q4.z * q4.z) + q4.w * q4.w);
      const float qd = fmaxf(qlen, 0.000001f);
      const float q[4] = {q4.x / qd, q4.y / qd, q4.z / qd, q4.w / qd};
      const float sx = hdsr_expf(kPacked ? r1.w : F.logscale[id * 3]) * F.splat_scale;
      const float sy = hdsr_expf(kPacked ? r2.x : F.logscale[id * 3 + 1]) * F.splat_scale;
      const float sz = hdsr_expf(kPacked ? r2.y : F.logscale[id * 3 + 2]) * F.splat_scale;
      const F3 a0 = mul(rotate_axis(q, f3(1.0f, 0.0f, 0.0f)), sx);
      const F3 a1 = mul(rotate_axis(q, f3(0.0f, 1.0f, 0.0f)), sy);
      const F3 a2 = mul(rotate_axis(q, f3(0.0f, 0.0f, 1.0f)), sz);
      const float len0 = dot(a0, a0), len1 = dot(a1, a1), len2 = dot(a2, a2);
      const float min_len = fminf(len0, fminf(len1, len2));
      const F3 normal_axis = (len0 == min_len) ? a0 : ((len1 == min_len) ? a1 : a2);

      const float* V = F.view;
      const F3 w0 = f3(V[0], V[4], V[8]);
      const F3 w1 = f3(V[1], V[5], V[9]);
      const F3 w2 = f3(V[2], V[6], V[10]);
      float vc[4];
      mat4_point(V, center, vc);
      const float vz = fminf(vc[2], -0.001f);
      const float j1x = F.fx / vz, j1y = F.fy / vz;
      const float j2x = (-j1x / vz) * vc[0];
      const float j2y = (-j1y / vz) * vc[1];
      const F3 tt0 = add(smul(j1x, w0), smul(j2x, w2));
      const F3 tt1 = add(smul(j1y, w1), smul(j2y, w2));
      const F3 b0 = f3(dot(a0, tt0), dot(a1, tt0), dot(a2, tt0));
      const F3 b1 = f3(dot(a0, tt1), dot(a1, tt1), dot(a2, tt1));
      const float cxx_raw = dot(b0, b0), cxy = dot(b0, b1), cyy_raw = dot(b1, b1);
      const float det_orig = cxx_raw * cyy_raw - cxy * cxy;
      const float cxx = cxx_raw + HDSR_LOW_PASS;
      const float cyy = cyy_raw + HDSR_LOW_PASS;
      const float det_blur = cxx * cyy - cxy * cxy;
      const float aa = sqrtf(fmaxf(det_orig / fmaxf(det_blur, 0.000001f), 0.0f));
      const float mid = 0.5f * (cxx + cyy);
      const float det = fmaxf(det_blur, 0.1f);
      const float lambda = mid + sqrtf(fmaxf(mid * mid - det, 0.1f));
      const float radius = ceilf(HDSR_SIGMA_RADIUS * sqrtf(lambda));
      if (!(radius <= 0.0f)) {
        const float det_inv = 1.0f / fmaxf(det_blur, 0.000001f);
        const float k0 = cyy * det_inv, k1 = -cxy * det_inv, k2 = cxx * det_inv;
        // differentiable mode: the opacity input is a logit (sigmoid evaluated here)
        const float o_raw = kPacked ? r2.z : F.opacity[id];
        const float o_in = F.diff ? 1.0f / (1.0f + hdsr_expf(-o_raw)) : o_raw;
        const float opac = fminf(fmaxf(o_in * aa, 0.0f), 0.99f);

        // tile box (gpu_project_splats.wgsl:210-218,229-232)
        const float ts = fmaxf(F.tile, 1.0f);
        const uint32_t mtx = max(F.tiles_x, 1u) - 1u, mty = max(F.tiles_y, 1u) - 1u;
        const float rc = fminf(radius, ts * 8.0f);
        const float minx = fmaxf(px - rc, 0.0f), miny = fmaxf(py - rc, 0.0f);
        const float maxx = fminf(px + rc, F.vw), maxy = fminf(py + rc, F.vh);
        const uint32_t t0x = min(hdsr_u32sat(floorf(minx / ts)), mtx);
        const uint32_t t0y = min(hdsr_u32sat(floorf(miny / ts)), mty);
        const uint32_t t1x = min(hdsr_u32sat(floorf(fmaxf((maxx - HDSR_FOOTPRINT_EPS) / ts, 0.0f))), mtx);
        const uint32_t t1y = min(hdsr_u32sat(floorf(fmaxf((maxy - HDSR_FOOTPRINT_EPS) / ts, 0.0f))), mty);
        bounds = (t0x & 0xFFu) | ((t0y & 0xFFu) << 8) | ((t1x & 0xFFu) << 16) | ((t1y & 0xFFu) << 24);
        const uint32_t bx0 = bounds & 0xFFu, by0 = (bounds >> 8) & 0xFFu;
        const uint32_t bx1 = (bounds >> 16) & 0xFFu, by1 = bounds >> 24;
        const uint32_t span = (bx1 - bx0 + 1u) * (by1 - by0 + 1u);  // u32 wrap as the WGSL
        const uint32_t ntiles = (bx1 >= bx0 && by1 >= by0) ? (bx1 - bx0 + 1u) * (by1 - by0 + 1u) : 0u;
        vis = 1;
        // 8 x 8-tile cells the box touches (the first level of the tile-list build)
        const uint32_t ncells =
            ntiles ? ((bx1 >> 3) - (bx0 >> 3) + 1u) * ((by1 >> 3) - (by0 >> 3) + 1u) : 0u;
        if (span > HDSR_LARGE_SPLAT_TILES) {
          is_large = 1;
          n_large = ntiles;
          c_large = ncells;
        } else {
          n_small = ntiles;
          c_small = ncells;
        }
        depth_bits = hdsr_f2u(depth);

        // normals (gpu_project_splats.wgsl:234-291)
        F3 nrm = normalize(normal_axis);
        const uint32_t nb = id * 3u;
        if (nb + 2u < F.normals_len) {
          const F3 baked = f3(F.normals[nb], F.normals[nb + 1], F.normals[nb + 2]);
          const float bl = vlen(baked);
          if (bl > 0.001f) nrm = vdiv(baked, bl);
        }
        const uint32_t db = F.n * 3u + id * 3u;
        if (db + 2u < F.normals_len) {
          const F3 detail = f3(F.normals[db], F.normals[db + 1], F.normals[db + 2]);
          const float dl = vlen(detail);
          if (dl > 0.001f) {
            F3 tc;
            if (len0 >= len1 && len0 >= len2) tc = a0;
            else if (len1 >= len2) tc = a1;
            else tc = a2;
            const F3 tp = sub(tc, mul(nrm, dot(tc, nrm)));
            const float tl = vlen(tp);
            if (tl > 0.001f) {
              const F3 T = vdiv(tp, tl);
              const F3 B = cross(nrm, T);
              const F3 dn = vdiv(detail, dl);
              nrm = normalize(add(add(mul(T, dn.x), mul(B, dn.y)), mul(nrm, dn.z)));
            }
          }
        }
        const F3 cam = f3(F.cam[0], F.cam[1], F.cam[2]);
        const F3 view_dir = normalize(sub(cam, center));
        const F3 nn = normalize(nrm);
        const F3 vv = normalize(view_dir);
        const F3 faced = dot(nn, vv) < 0.0f ? f3(-nn.x, -nn.y, -nn.z) : nn;
        const float conf = fminf(fmaxf(dot(faced, view_dir), 0.0f), 1.0f);
        const float* M = F.nm;
        F3 wn;
        wn.x = (M[0] * faced.x + M[3] * faced.y) + M[6] * faced.z;
        wn.y = (M[1] * faced.x + M[4] * faced.y) + M[7] * faced.z;
        wn.z = (M[2] * faced.x + M[5] * faced.y) + M[8] * faced.z;
        float ox, oy;
        oct_encode(normalize(wn), &ox, &oy);

        // spherical harmonics (gpu_project_splats.wgsl:293-351)
        const uint32_t deg = F.sh_degree;
        const float* shp = kPacked ? nullptr : F.shdata + (size_t)id * F.sh_stride;
        // SH row float k: from the registers (packed) or the caller's array
        auto S = [&](int k) -> float {
          if (kPacked) {
            const float4 v = rv[3 + (k >> 2)];
            return (k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w;
          }
          return shp[k];
        };
        float c[3] = {S(0), S(1), S(2)};
        if (deg >= 1u) {
          const float x = view_dir.x, y = view_dir.y, z = view_dir.z;
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) c[ch] = c[ch] + SH_C1 * ((S(3 + ch) * y + S(6 + ch) * z) + S(9 + ch) * x);
          if (deg >= 2u) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            const float b2[5] = {SH_C2_0 * xy, SH_C2_1 * yz, SH_C2_2 * ((2.0f * zz - xx) - yy), SH_C2_3 * xz,
                                 SH_C2_4 * (xx - yy)};
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
              float acc = S(12 + ch) * b2[0];
#pragma unroll
              for (int j = 1; j < 5; ++j) acc = acc + S(12 + 3 * j + ch) * b2[j];
              c[ch] = c[ch] + acc;
            }
            if (deg >= 3u) {
              const float b3[7] = {(SH_C3_0 * y) * (3.0f * xx - yy),
                                   ((SH_C3_1 * x) * y) * z,
                                   (SH_C3_2 * y) * ((4.0f * zz - xx) - yy),
                                   (SH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy),
                                   (SH_C3_4 * x) * ((4.0f * zz - xx) - yy),
                                   (SH_C3_5 * z) * (xx - yy),
                                   (SH_C3_6 * x) * (xx - 3.0f * yy)};
#pragma unroll
              for (int ch = 0; ch < 3; ++ch) {
                float acc = S(27 + ch) * b3[0];
#pragma unroll
                for (int j = 1; j < 7; ++j) acc = acc + S(27 + 3 * j + ch) * b3[j];
                c[ch] = c[ch] + acc;
              }
            }
          }
        }
        const float cr = fminf(fmaxf(c[0], 0.0f), 1.0f);
        const float cg = fminf(fmaxf(c[1], 0.0f), 1.0f);
        const float cb = fminf(fmaxf(c[2], 0.0f), 1.0f);
        // emissive: the last three floats of the SH row (stride 3 (deg + 1)^2 + 3: offset 3, 12, 27 or 48)
        const uint32_t eo = F.sh_stride - 3u;
        float em[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          em[i] = kPacked ? (eo == 3u ? S(3 + i) : eo == 12u ? S(12 + i) : eo == 27u ? S(27 + i) : S(48 + i))
                          : shp[eo + i];

        const uint32_t op16 = hdsr_f16_bits(opac);
        h0 = make_uint4(hdsr_f2u(px), hdsr_f2u(py), hdsr_f2u(k0), hdsr_f2u(k1));
        h1 = make_uint4(hdsr_f2u(k2), depth_bits, hdsr_pack2x16(cr, cg), hdsr_f16_bits(cb) | (op16 << 16));
        c0 = make_uint4(kPacked ? hdsr_f2u(r2.w) : F.material[id], hdsr_f16_bits(radius) | (op16 << 16), hdsr_pack2x16(ox, oy),
                        hdsr_pack2x16(conf, 0.0f));
        c1 = make_uint4(hdsr_pack2x16(em[0], em[1]), hdsr_pack2x16(em[2], 0.0f), 0u, 0u);
        if (F.diff && b.drec) {
          // fp32 record for the differentiable rasterizer: no f16 rounding of colour / opacity
          float4* d = b.drec + (size_t)g * 3;
          d[0] = make_float4(px, py, k0, k1);
          d[1] = make_float4(k2, opac, cr, cg);
          d[2] = make_float4(cb, depth, 0.0f, 0.0f);
        }
      }
    }
  }
  b.hot[(size_t)g * 2] = h0;
  b.hot[(size_t)g * 2 + 1] = h1;
  if (b.write_cold) {  // uniform: G-buffer or debug readback only
    b.cold[(size_t)g * 2] = c0;
    b.cold[(size_t)g * 2 + 1] = c1;
  }
  b.bounds[g] = bounds;
  b.depth[g] = depth_bits;
  // ln(max(depth, 1e-6)) once per splat; the per-tile keys and buckets subtract the tile's
  // logMin from it (hdsr_log_pos evaluates the same expression)
  // 8-byte records the two tile-list walks gather per (cell, splat) pair: one line per pair, 16
  // splat ranks per 128-byte line
  b.bdep[g] = make_uint2(bounds, depth_bits);
  b.bln[g] = make_uint2(bounds, hdsr_f2u(hdsr_logf(fmaxf(hdsr_u2f(depth_bits), 1e-6f))));

  // workgroup reduction of the six counts
  __shared__ uint32_t red[4][6];
  const uint32_t s0 = wave_sum(n_small), s1 = wave_sum(n_large), s2 = wave_sum(is_large), s3 = wave_sum(vis);
  const uint32_t s4 = wave_sum(c_small), s5 = wave_sum(c_large);
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    red[w][0] = s0;
    red[w][1] = s1;
    red[w][2] = s2;
    red[w][3] = s3;
    red[w][4] = s4;
    red[w][5] = s5;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 t = make_uint4(0, 0, 0, 0), c = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      t.x += red[i][0];
      t.y += red[i][1];
      t.z += red[i][2];
      t.w += red[i][3];
      c.x += red[i][4];
      c.y += red[i][5];
    }
    b.block_cnt[blk] = t;
    b.block_cc[blk] = c;
  }
}

// Rebuild of the 14-word reference record for frame f (debug readback only).
__global__ void k_cache14(Bufs b, uint32_t f, uint32_t n, uint32_t* dst) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t g = b.frames[f].splat_base + r;
  uint32_t* w = dst + (size_t)r * 14;
  const uint32_t bd = b.bounds[g];
  if (bd == HDSR_SENTINEL) {
    for (int i = 0; i < 14; ++i) w[i] = 0;
    w[7] = HDSR_SENTINEL;
    return;
  }
  const uint4 h0 = b.hot[(size_t)g * 2], h1 = b.hot[(size_t)g * 2 + 1];
  const uint4 c0 = b.cold[(size_t)g * 2], c1 = b.cold[(size_t)g * 2 + 1];
  w[0] = h0.x;
  w[1] = h0.y;
  w[2] = h0.z;
  w[3] = h0.w;
  w[4] = h1.x;
  w[5] = c0.x;
  w[6] = c0.y;
  w[7] = bd;
  w[8] = c0.z;
  w[9] = h1.z;
  w[10] = h1.w & 0xFFFFu;
  w[11] = c1.x;
  w[12] = c1.y;
  w[13] = c0.w;
}

// One thread per float4 slot of the packed records (coalesced writes; runs once per bound scene).
__global__ void k_pack_scene(const float* __restrict__ pos, const float* __restrict__ logscale,
                             const float* __restrict__ rot, const float* __restrict__ opacity,
                             const uint32_t* __restrict__ material, const float* __restrict__ shdata, uint32_t n,
                             uint32_t sh_stride, uint32_t s4, uint64_t total, float4* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const uint32_t i = (uint32_t)(q / s4), slot = (uint32_t)(q % s4);
  float v[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t f = slot * 4 + k;
    float x = 0.0f;
    if (f < 4) x = rot[(size_t)i * 4 + f];
    else if (f < 7) x = pos[(size_t)i * 3 + (f - 4)];
    else if (f < 10) x = logscale[(size_t)i * 3 + (f - 7)];
    else if (f == 10) x = opacity[i];
    else if (f == 11) x = hdsr_u2f(material[i]);
    else if (f - 12 < sh_stride) x = shdata[(size_t)i * sh_stride + (f - 12)];
    v[k] = x;
  }
  out[q] = make_float4(v[0], v[1], v[2], v[3]);
}

}  // namespace

void launch_project(const Bufs& b, hipStream_t stream) {
  if (!b.splat_blocks) return;
  if (b.packed)
    hipLaunchKernelGGL(k_project<true>, dim3(b.splat_blocks), dim3(kSplatBlock), 0, stream, b);
  else
    hipLaunchKernelGGL(k_project<false>, dim3(b.splat_blocks), dim3(kSplatBlock), 0, stream, b);
}

void launch_pack_scene(const float* pos, const float* logscale, const float* rot, const float* opacity,
                       const uint32_t* material, const float* shdata, uint32_t n, uint32_t sh_stride, float4* out,
                       hipStream_t stream) {
  if (!n) return;
  const uint32_t s4 = packed_stride4(sh_stride);
  const uint64_t total = (uint64_t)n * s4;
  hipLaunchKernelGGL(k_pack_scene, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, pos, logscale, rot,
                     opacity, material, shdata, n, sh_stride, s4, total, out);
}

void launch_cache14(const Bufs& b, uint32_t frame, uint32_t n, uint32_t* dst, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_cache14, dim3((n + 255) / 256), dim3(256), 0, stream, b, frame, n, dst);
}

}  // namespace hdsr
