// k_diff.hip - backward of the differentiable mode: rasterizer, then projection.
//
// The forward of this mode is the plain path with two changes: the opacity input is a logit
// (sigmoid applied in the projection), and the rasterizer reads an fp32 record (drec) instead
// of the f16-packed colour / opacity of the reference cache.  The per-tile lists and their order
// are those of the plain path and are held fixed (discrete) in the backward.
//
// Rasterizer backward (one 256-thread workgroup per 16x16 tile, one pixel per thread).  The
// forward kept, per pixel, the final transmittance T_f and one past the list index of the last
// entry that blended.  With w_i = a_i T_i, C = sum_i c_i w_i + T_f * bg and A = 1 - T_f:
//   dC/dc_i = w_i
//   dC/da_i = c_i T_i - S_i / (1 - a_i),   S_i = sum_{j>i} c_j w_j + T_f * bg
//   dA/da_i = T_f / (1 - a_i)
// so walking each pixel's list back to front from its last entry, with T_i = T_{i+1} / (1 - a_i)
// recovered by division and S accumulated, gives every entry's gradient in one pass.  Through
// a = min(0.99, o * exp(p)) (zero where the clamp is active) and
// p = -(k0 dx^2 + 2 k1 dx dy + k2 dy^2) / 2 with d = pixel - centre, the 9 gradient slots per
// splat are (centre x, y, k0, k1, k2, opacity, r, g, b).  Per entry, each wave sums its 64 lanes
// with cross-lane adds, the four waves' sums meet in LDS, and one global atomic per
// (splat, slot) adds the tile's total into gacc.
//
// Projection backward (one thread per sort rank): recomputes the forward values in fp32 and
// applies the chain rule through the WGSL projection (gpu_project_splats.wgsl:112-358): screen
// position, the EWA covariance with the 0.3 low-pass and the sqrt(det / det_blur) opacity
// factor, the conic, the sigmoid of the opacity logit, the SH colour (all-positive basis, view
// direction camera - centre) and its [0, 1] clamp.
#include <algorithm>

#include "hdsr_internal.h"
#include "proj_common.h"
#include "raster_common.h"

namespace hdsr {
namespace {

constexpr uint32_t kBatch = 256;
constexpr int kSlots = 9;

__global__ void k_fill_zero(float* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Sums each of the 9 slots over the wave's 64 lanes with 13 cross-lane adds instead of 9 x 6: at
// each butterfly step a lane keeps half of its remaining slots and sends the other half to its
// partner (9 -> 5 -> 3 -> 2 -> 1 slots over the lane bits 5, 4, 3, 2; bits 1, 0 finish the last
// one).  Returns the wave total of slot_of_lane(lane) (see below) on every lane.
__device__ __forceinline__ float wave_sum9(const float (&v)[kSlots], uint32_t lane) {
  const bool b5 = lane & 32u, b4 = lane & 16u, b3 = lane & 8u, b2 = lane & 4u;
  float a[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {  // keep 0..4 (bit 5 clear) or 5..8 (set)
    const float send = b5 ? v[k] : (k < 4 ? v[5 + k] : 0.0f);
    const float recv = __shfl_xor(send, 32, 64);
    a[k] = b5 ? (k < 4 ? v[5 + k] + recv : 0.0f) : v[k] + recv;
  }
  float c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // keep a0..a2 or a3, a4
    const float send = b4 ? a[k] : (k < 2 ? a[3 + k] : 0.0f);
    const float recv = __shfl_xor(send, 16, 64);
    c[k] = b4 ? (k < 2 ? a[3 + k] + recv : 0.0f) : a[k] + recv;
  }
  float d[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {  // keep c0, c1 or c2
    const float send = b3 ? c[k] : (k < 1 ? c[2 + k] : 0.0f);
    const float recv = __shfl_xor(send, 8, 64);
    d[k] = b3 ? (k < 1 ? c[2 + k] + recv : 0.0f) : c[k] + recv;
  }
  float e = (b2 ? d[1] : d[0]) + __shfl_xor(b2 ? d[0] : d[1], 4, 64);
  e += __shfl_xor(e, 2, 64);
  e += __shfl_xor(e, 1, 64);
  return e;
}

// The slot whose total wave_sum9 leaves on 'lane', or -1 (a padding position).
__device__ __forceinline__ int slot_of_lane(uint32_t lane) {
  const int d = (lane & 4u) ? 1 : 0;
  int c = d;
  if (lane & 8u) c = d ? -1 : 2;
  int a = c;
  if (lane & 16u) a = (c < 0 || c > 1) ? -1 : 3 + c;
  if (a < 0) return -1;
  if (lane & 32u) return a > 3 ? -1 : 5 + a;
  return a;
}

__global__ __launch_bounds__(256) void k_raster_bwd(Bufs b, const float* __restrict__ grad) {
  __shared__ float4 s_a[kBatch];  // cx, cy, -k0/2, -k1
  __shared__ float4 s_b[kBatch];  // -k2/2, opacity, r, g
  __shared__ float4 s_c[kBatch];  // b, 0, reach y0, reach y1
  __shared__ uint32_t s_idx[kBatch];
  __shared__ uint32_t s_gid[kBatch];
  __shared__ float s_g[4][kSlots][kBatch];
  __shared__ uint32_t s_keep[4];
  __shared__ uint32_t s_top[4];

  const FrameDesc& F = b.frames[0];
  const uint32_t t = blockIdx.x;
  const uint32_t tx = t % F.tiles_x, ty = t / F.tiles_x;
  const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
  const uint32_t start = b.tile_offsets[m];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t lx = tid & 15u, ly = tid >> 4;
  const uint32_t X = tx * 16 + lx, Y = ty * 16 + ly;
  const bool inside = X < F.width && Y < F.height;
  const size_t o = inside ? (size_t)Y * F.width + X : 0;
  const uint32_t my_last = inside ? F.last_idx[o] : 0u;
  const float Tf = inside ? F.t_final[o] : 1.0f;
  const float4 gin = inside ? reinterpret_cast<const float4*>(grad)[o] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);

  const uint32_t wm = wave_max(my_last);  // no entry at or past this index blends in this wave
  if (lane == 0) s_top[wave] = wm;
  const int my_slot = (lane & 3u) == 0u ? slot_of_lane(lane) : -1;
  __syncthreads();
  const uint32_t top = max(max(s_top[0], s_top[1]), max(s_top[2], s_top[3]));

  const float p00x = (float)(X & ~1u) + 0.5f, p00y = (float)(Y & ~1u) + 0.5f;
  const bool odd_x = (lx & 1u) != 0, odd_y = (ly & 1u) != 0;
  const float tile_x0 = (float)(tx * 16) + 0.5f, tile_y0 = (float)(ty * 16) + 0.5f;
  const float band_y0 = tile_y0 + (float)(wave * 4), band_y1 = band_y0 + 3.0f;

  float T = Tf;
  float Sr = Tf * 0.02f, Sg = Tf * 0.02f, Sb = Tf * 0.04f;
  const float gAT = gin.w * Tf;

  for (int32_t base = top ? (int32_t)((top - 1u) / kBatch * kBatch) : -1; base >= 0; base -= (int32_t)kBatch) {
    const uint32_t i = (uint32_t)base + tid;
    bool keep = false;
    float4 r0 = make_float4(0, 0, 0, 0), r1 = r0, r2 = r0;
    float2 reach = make_float2(0.0f, 0.0f);
    uint32_t g = 0;
    if (i < top) {
      g = b.entries[start + i];
      const float4* d = b.drec + (size_t)g * 3;
      r0 = d[0];
      r1 = d[1];
      r2 = d[2];
      reach = reach_y(r0.x, r0.y, r0.z, r0.w, r1.x, tile_x0, tile_x0 + 15.0f, tile_y0, tile_y0 + 15.0f);
      keep = !(reach.x > reach.y);
    }
    const unsigned long long km = __ballot(keep);
    if (lane == 0) s_keep[wave] = (uint32_t)__popcll(km);
    __syncthreads();
    uint32_t slot = (uint32_t)__popcll(km & ((1ull << lane) - 1ull));
    for (uint32_t q = 0; q < wave; ++q) slot += s_keep[q];
    const uint32_t nb = s_keep[0] + s_keep[1] + s_keep[2] + s_keep[3];
    if (keep) {
      s_a[slot] = make_float4(r0.x, r0.y, r0.z * -0.5f, -r0.w);
      s_b[slot] = make_float4(r1.x * -0.5f, r1.y, r1.z, r1.w);
      s_c[slot] = make_float4(r2.x, 0.0f, reach.x, reach.y);
      s_idx[slot] = i;
      s_gid[slot] = g;
    }
    __syncthreads();
    for (int32_t j = (int32_t)nb - 1; j >= 0; --j) {
      float gv[kSlots];
#pragma unroll
      for (int k = 0; k < kSlots; ++k) gv[k] = 0.0f;
      float tot = 0.0f;
      const float4 C = s_c[j];
      if (s_idx[j] < wm && !(C.z > band_y1 || C.w < band_y0)) {  // uniform per wave
        const float4 A = s_a[j];
        const float4 B = s_b[j];
        float dx = p00x - A.x, dy = p00y - A.y;
        dx = odd_x ? dx + 1.0f : dx;
        dy = odd_y ? dy + 1.0f : dy;
        const float pw = (A.z * dx * dx + A.w * dx * dy) + B.x * dy * dy;
        const bool act = s_idx[j] < my_last && pw > -4.0f;
        if (__any(act)) {  // uniform
          // the forward's gaussian and alpha, bit for bit (exp only consumed where power > -4)
          const float gauss = hdsr_expf_gauss(fminf(fmaxf(pw, -86.0f), 88.0f));
          const float alpha = fminf(0.99f, B.y * gauss);
          const bool valid = act && alpha > (1.0f / 255.0f);
          if (__any(valid)) {
            // every lane computes; lanes that did not blend contribute zeros and keep T, S
            const float om = 1.0f - alpha;
            const float Tj = T / om;
            const float w = alpha * Tj;
            const float dA = ((gin.x * (B.z * Tj - Sr / om) + gin.y * (B.w * Tj - Sg / om)) +
                              gin.z * (C.x * Tj - Sb / om)) + gAT / om;
            gv[6] = valid ? gin.x * w : 0.0f;
            gv[7] = valid ? gin.y * w : 0.0f;
            gv[8] = valid ? gin.z * w : 0.0f;
            // below the 0.99 clamp: d alpha / d power = o exp(power) = alpha, d alpha / d o = exp(power)
            const bool cm = valid && B.y * gauss < 0.99f;
            const float gp = cm ? dA * alpha : 0.0f;
            const float k0 = -2.0f * A.z, k1 = -A.w, k2 = -2.0f * B.x;
            gv[0] = gp * (k0 * dx + k1 * dy);
            gv[1] = gp * (k1 * dx + k2 * dy);
            gv[2] = -0.5f * gp * dx * dx;
            gv[3] = -gp * dx * dy;
            gv[4] = -0.5f * gp * dy * dy;
            gv[5] = cm ? dA * gauss : 0.0f;
            Sr = valid ? Sr + B.z * w : Sr;
            Sg = valid ? Sg + B.w * w : Sg;
            Sb = valid ? Sb + C.x * w : Sb;
            T = valid ? Tj : T;
            tot = wave_sum9(gv, lane);
          }
        }
      }
      if (my_slot >= 0) s_g[wave][my_slot][j] = tot;
    }
    __syncthreads();
    if (tid < nb) {
      float* dst = b.gacc + (size_t)s_gid[tid] * kSlots;
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        const float v = (s_g[0][k][tid] + s_g[1][k][tid]) + (s_g[2][k][tid] + s_g[3][k][tid]);
        if (v != 0.0f) unsafeAtomicAdd(dst + k, v);
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void atomic_add3(float* p, F3 v) {
  unsafeAtomicAdd(p, v.x);
  unsafeAtomicAdd(p + 1, v.y);
  unsafeAtomicAdd(p + 2, v.z);
}

__global__ __launch_bounds__(256) void k_project_bwd(Bufs b, DiffGrads out) {
  const FrameDesc& F = b.frames[0];
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= F.n) return;
  const uint32_t g = F.splat_base + r;
  if (b.bounds[g] == HDSR_SENTINEL) return;
  float ga[kSlots];
  bool any = false;
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    ga[k] = b.gacc[(size_t)g * kSlots + k];
    any |= ga[k] != 0.0f;
  }
  if (!any) return;
  const uint32_t id = F.sorted[r];
  if (id >= F.n) return;

  // ---- forward values (as k_project) ----
  const F3 center = f3(F.pos[id * 3], F.pos[id * 3 + 1], F.pos[id * 3 + 2]);
  float clip[4];
  mat4_point(F.vp, center, clip);
  const float safe_w = fmaxf(clip[3], HDSR_MIN_CLIP_W);
  const float ndc_x = clip[0] / safe_w, ndc_y = clip[1] / safe_w;
  const float4 q4 = *reinterpret_cast<const float4*>(F.rot + (size_t)id * 4);
  const float qlen = sqrtf(((q4.x * q4.x + q4.y * q4.y) + q4.z * q4.z) + q4.w * q4.w);
  const float qd = fmaxf(qlen, 0.000001f);
  const float q[4] = {q4.x / qd, q4.y / qd, q4.z / qd, q4.w / qd};
  const float s[3] = {hdsr_expf(F.logscale[id * 3]) * F.splat_scale, hdsr_expf(F.logscale[id * 3 + 1]) * F.splat_scale,
                      hdsr_expf(F.logscale[id * 3 + 2]) * F.splat_scale};
  const F3 e[3] = {f3(1.0f, 0.0f, 0.0f), f3(0.0f, 1.0f, 0.0f), f3(0.0f, 0.0f, 1.0f)};
  F3 rax[3], a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rax[i] = rotate_axis(q, e[i]);
    a[i] = mul(rax[i], s[i]);
  }
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
  const float b0[3] = {dot(a[0], tt0), dot(a[1], tt0), dot(a[2], tt0)};
  const float b1[3] = {dot(a[0], tt1), dot(a[1], tt1), dot(a[2], tt1)};
  const float cxx_raw = (b0[0] * b0[0] + b0[1] * b0[1]) + b0[2] * b0[2];
  const float cxy = (b0[0] * b1[0] + b0[1] * b1[1]) + b0[2] * b1[2];
  const float cyy_raw = (b1[0] * b1[0] + b1[1] * b1[1]) + b1[2] * b1[2];
  const float det_orig = cxx_raw * cyy_raw - cxy * cxy;
  const float cxx = cxx_raw + HDSR_LOW_PASS;
  const float cyy = cyy_raw + HDSR_LOW_PASS;
  const float det_blur = cxx * cyy - cxy * cxy;
  const float dbc = fmaxf(det_blur, 0.000001f);
  const float ratio = det_orig / dbc;
  const float aa = sqrtf(fmaxf(ratio, 0.0f));
  const float det_inv = 1.0f / dbc;
  const float logit = F.opacity[id];
  const float sig = 1.0f / (1.0f + hdsr_expf(-logit));
  const float oa = sig * aa;

  F3 gm = f3(0.0f, 0.0f, 0.0f);

  // ---- colour: SH coefficients and the view direction ----
  {
    const F3 cam = f3(F.cam[0], F.cam[1], F.cam[2]);
    const F3 v = sub(cam, center);
    const float vl = vlen(v);
    const F3 dir = vdiv(v, vl);
    const float x = dir.x, y = dir.y, z = dir.z;
    const uint32_t deg = F.sh_degree;
    const float* sh = F.shdata + (size_t)id * F.sh_stride;
    float* gsh = out.sh_data + (size_t)id * F.sh_stride;
    // forward colour (unclamped) for the clamp masks
    float c[3] = {sh[0], sh[1], sh[2]};
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    const float b1v[3] = {SH_C1 * y, SH_C1 * z, SH_C1 * x};
    const float b2[5] = {SH_C2_0 * xy, SH_C2_1 * yz, SH_C2_2 * ((2.0f * zz - xx) - yy), SH_C2_3 * xz,
                         SH_C2_4 * (xx - yy)};
    const float b3[7] = {(SH_C3_0 * y) * (3.0f * xx - yy),
                         ((SH_C3_1 * x) * y) * z,
                         (SH_C3_2 * y) * ((4.0f * zz - xx) - yy),
                         (SH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy),
                         (SH_C3_4 * x) * ((4.0f * zz - xx) - yy),
                         (SH_C3_5 * z) * (xx - yy),
                         (SH_C3_6 * x) * (xx - 3.0f * yy)};
    if (deg >= 1u) {
      const float* s1 = sh + 3;
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) c[ch] = c[ch] + SH_C1 * ((s1[ch] * y + s1[3 + ch] * z) + s1[6 + ch] * x);
      if (deg >= 2u) {
        const float* s2 = s1 + 9;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          float acc = s2[ch] * b2[0];
#pragma unroll
          for (int j = 1; j < 5; ++j) acc = acc + s2[3 * j + ch] * b2[j];
          c[ch] = c[ch] + acc;
        }
        if (deg >= 3u) {
          const float* s3 = s2 + 15;
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            float acc = s3[ch] * b3[0];
#pragma unroll
            for (int j = 1; j < 7; ++j) acc = acc + s3[3 * j + ch] * b3[j];
            c[ch] = c[ch] + acc;
          }
        }
      }
    }
    float gc[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) gc[ch] = (c[ch] > 0.0f && c[ch] < 1.0f) ? ga[6 + ch] : 0.0f;
    if (gc[0] != 0.0f || gc[1] != 0.0f || gc[2] != 0.0f) {
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) unsafeAtomicAdd(gsh + ch, gc[ch]);
      F3 gd = f3(0.0f, 0.0f, 0.0f);
      if (deg >= 1u) {
        const float* s1 = sh + 3;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) unsafeAtomicAdd(gsh + 3 + 3 * j + ch, b1v[j] * gc[ch]);
        // d(colour)/d(dir) of band 1: C1 * (s[0] y + s[1] z + s[2] x)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          gd.x += gc[ch] * SH_C1 * s1[6 + ch];
          gd.y += gc[ch] * SH_C1 * s1[ch];
          gd.z += gc[ch] * SH_C1 * s1[3 + ch];
        }
        if (deg >= 2u) {
          const float* s2 = s1 + 9;
          // derivatives of the band-2 basis along x, y, z
          const float dbx[5] = {SH_C2_0 * y, 0.0f, SH_C2_2 * (-2.0f * x), SH_C2_3 * z, SH_C2_4 * (2.0f * x)};
          const float dby[5] = {SH_C2_0 * x, SH_C2_1 * z, SH_C2_2 * (-2.0f * y), 0.0f, SH_C2_4 * (-2.0f * y)};
          const float dbz[5] = {0.0f, SH_C2_1 * y, SH_C2_2 * (4.0f * z), SH_C2_3 * x, 0.0f};
#pragma unroll
          for (int j = 0; j < 5; ++j)
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
              unsafeAtomicAdd(gsh + 12 + 3 * j + ch, b2[j] * gc[ch]);
              const float sc = gc[ch] * s2[3 * j + ch];
              gd.x += sc * dbx[j];
              gd.y += sc * dby[j];
              gd.z += sc * dbz[j];
            }
          if (deg >= 3u) {
            const float* s3 = s2 + 15;
            const float dbx3[7] = {SH_C3_0 * (6.0f * xy), SH_C3_1 * yz, SH_C3_2 * (-2.0f * xy), SH_C3_3 * (-6.0f * xz),
                                   SH_C3_4 * ((4.0f * zz - 3.0f * xx) - yy), SH_C3_5 * (2.0f * xz),
                                   SH_C3_6 * (3.0f * xx - 3.0f * yy)};
            const float dby3[7] = {SH_C3_0 * (3.0f * xx - 3.0f * yy), SH_C3_1 * xz,
                                   SH_C3_2 * ((4.0f * zz - xx) - 3.0f * yy), SH_C3_3 * (-6.0f * yz),
                                   SH_C3_4 * (-2.0f * xy), SH_C3_5 * (-2.0f * yz), SH_C3_6 * (-6.0f * xy)};
            const float dbz3[7] = {0.0f, SH_C3_1 * xy, SH_C3_2 * (8.0f * yz),
                                   SH_C3_3 * ((6.0f * zz - 3.0f * xx) - 3.0f * yy), SH_C3_4 * (8.0f * xz),
                                   SH_C3_5 * (xx - yy), 0.0f};
#pragma unroll
            for (int j = 0; j < 7; ++j)
#pragma unroll
              for (int ch = 0; ch < 3; ++ch) {
                unsafeAtomicAdd(gsh + 27 + 3 * j + ch, b3[j] * gc[ch]);
                const float sc = gc[ch] * s3[3 * j + ch];
                gd.x += sc * dbx3[j];
                gd.y += sc * dby3[j];
                gd.z += sc * dbz3[j];
              }
          }
        }
        // dir = v / |v|, v = cam - centre
        const float pr = dot(dir, gd);
        const F3 gv = vdiv(sub(gd, mul(dir, pr)), vl);
        gm = sub(gm, gv);
      }
    }
  }

  // ---- opacity ----
  const float g_oa = (oa > 0.0f && oa < 0.99f) ? ga[5] : 0.0f;
  const float g_aa = g_oa * sig;
  unsafeAtomicAdd(out.opacity_logits + id, (g_oa * aa) * (sig * (1.0f - sig)));

  // ---- conic and opacity factor -> 2D covariance ----
  const float gk0 = ga[2], gk1 = ga[3], gk2 = ga[4];
  float g_cxx = 0.0f, g_cyy = 0.0f, g_cxy = 0.0f;  // w.r.t. the low-passed cxx, cyy and cxy
  float g_cxx_raw = 0.0f, g_cyy_raw = 0.0f;
  g_cyy += gk0 * det_inv;
  g_cxy += -gk1 * det_inv;
  g_cxx += gk2 * det_inv;
  float g_db = 0.0f;
  if (det_blur > 0.000001f) {
    const float g_dinv = (gk0 * cyy - gk1 * cxy) + gk2 * cxx;
    g_db += -g_dinv * det_inv * det_inv;
  }
  if (ratio > 0.0f && g_aa != 0.0f) {
    const float g_ratio = g_aa * 0.5f / aa;
    const float g_do = g_ratio / dbc;
    if (det_blur > 0.000001f) g_db += -g_ratio * ratio / dbc;
    g_cxx_raw += g_do * cyy_raw;
    g_cyy_raw += g_do * cxx_raw;
    g_cxy += -2.0f * cxy * g_do;
  }
  g_cxx += g_db * cyy;
  g_cyy += g_db * cxx;
  g_cxy += -2.0f * cxy * g_db;
  g_cxx_raw += g_cxx;
  g_cyy_raw += g_cyy;

  // ---- covariance = (M^T tt)^T (M^T tt), M = [a0 a1 a2] ----
  float gb0[3], gb1[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gb0[i] = 2.0f * g_cxx_raw * b0[i] + g_cxy * b1[i];
    gb1[i] = 2.0f * g_cyy_raw * b1[i] + g_cxy * b0[i];
  }
  F3 g_tt0 = f3(0.0f, 0.0f, 0.0f), g_tt1 = f3(0.0f, 0.0f, 0.0f);
  F3 g_u = f3(0.0f, 0.0f, 0.0f);
  float g_w = 0.0f;
  const F3 u = f3(q[1], q[2], q[3]);
  float gls[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const F3 g_a = add(smul(gb0[i], tt0), smul(gb1[i], tt1));
    g_tt0 = add(g_tt0, smul(gb0[i], a[i]));
    g_tt1 = add(g_tt1, smul(gb1[i], a[i]));
    gls[i] = dot(g_a, rax[i]) * s[i];  // s = exp(log-scale) * splat_scale
    // rotate_axis: r = e + 2 u x h, h = u x e + w e
    const F3 g_r2 = smul(2.0f * s[i], g_a);
    const F3 h = add(cross(u, e[i]), smul(q[0], e[i]));
    g_u = add(g_u, cross(h, g_r2));
    const F3 g_h = cross(g_r2, u);
    g_u = add(g_u, cross(e[i], g_h));
    g_w += dot(g_h, e[i]);
  }
  unsafeAtomicAdd(out.log_scales + (size_t)id * 3, gls[0]);
  unsafeAtomicAdd(out.log_scales + (size_t)id * 3 + 1, gls[1]);
  unsafeAtomicAdd(out.log_scales + (size_t)id * 3 + 2, gls[2]);
  {
    // q_n = q / max(|q|, 1e-6)
    const float gq[4] = {g_w, g_u.x, g_u.y, g_u.z};
    float* dst = out.rotations + (size_t)id * 4;
    if (qlen > 0.000001f) {
      const float pr = ((q[0] * gq[0] + q[1] * gq[1]) + q[2] * gq[2]) + q[3] * gq[3];
#pragma unroll
      for (int k = 0; k < 4; ++k) unsafeAtomicAdd(dst + k, (gq[k] - q[k] * pr) / qlen);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) unsafeAtomicAdd(dst + k, gq[k] / qd);
    }
  }

  // ---- projection Jacobian -> view-space centre ----
  {
    const float g_j1x = dot(g_tt0, w0), g_j2x = dot(g_tt0, w2);
    const float g_j1y = dot(g_tt1, w1), g_j2y = dot(g_tt1, w2);
    const float G1x = g_j1x + g_j2x * (-vc[0] / vz);
    const float G1y = g_j1y + g_j2y * (-vc[1] / vz);
    const float g_vc0 = g_j2x * (-j1x / vz);
    const float g_vc1 = g_j2y * (-j1y / vz);
    const float g_vz = (G1x * (-j1x / vz) + g_j2x * (j1x * vc[0] / (vz * vz))) +
                       (G1y * (-j1y / vz) + g_j2y * (j1y * vc[1] / (vz * vz)));
    const float g_vc2 = vc[2] < -0.001f ? g_vz : 0.0f;
    gm.x += (V[0] * g_vc0 + V[1] * g_vc1) + V[2] * g_vc2;
    gm.y += (V[4] * g_vc0 + V[5] * g_vc1) + V[6] * g_vc2;
    gm.z += (V[8] * g_vc0 + V[9] * g_vc1) + V[10] * g_vc2;
  }
  // ---- screen position ----
  {
    const float g_nx = ga[0] * 0.5f * F.vw;
    const float g_ny = -ga[1] * 0.5f * F.vh;
    const float g_c0 = g_nx / safe_w, g_c1 = g_ny / safe_w;
    const float g_c3 = clip[3] > HDSR_MIN_CLIP_W ? -(g_nx * ndc_x + g_ny * ndc_y) / safe_w : 0.0f;
    const float* P = F.vp;
    gm.x += (P[0] * g_c0 + P[1] * g_c1) + P[3] * g_c3;
    gm.y += (P[4] * g_c0 + P[5] * g_c1) + P[7] * g_c3;
    gm.z += (P[8] * g_c0 + P[9] * g_c1) + P[11] * g_c3;
  }
  atomic_add3(out.means + (size_t)id * 3, gm);
}

}  // namespace

void launch_backward(const Bufs& b, uint32_t n, const float* grad_rgba, const DiffGrads& out, hipStream_t stream) {
  const size_t ng = (size_t)b.splat_blocks * kSplatBlock * kSlots;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<size_t>((ng + 255) / 256, 4096)), dim3(256), 0, stream,
                     b.gacc, ng);
  if (b.total_tiles) hipLaunchKernelGGL(k_raster_bwd, dim3(b.total_tiles), dim3(256), 0, stream, b, grad_rgba);
  if (n) hipLaunchKernelGGL(k_project_bwd, dim3((n + 255) / 256), dim3(256), 0, stream, b, out);
}

}  // namespace hdsr
