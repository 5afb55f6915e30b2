// k_deferred.hip - the first consumers of the G-buffer: screen-space normal reconstruction and
// the depth-guided bilateral normal filter.
//
// Reference: main of src/shaders/gpu_screen_space_normals.wgsl:59-120 and main of
// src/shaders/bilateral_normal_filter.wgsl:39-86, encoded by src/splatRenderer.ts:1042-1069
// (parameters :350-363 and :417-423).  Inputs are the compositor's fp32 depth and oct-encoded
// normal words (pack2x16float of octEncode); both passes write oct-encoded normal words.
//
// Arithmetic is IEEE fp32 without contraction in the WGSL's operation order: a matrix-vector
// product sums its four column terms left to right, normalize(v) = v / sqrt(dot(v, v)),
// mix(a, b, t) = a * (1 - t) + b * t, smoothstep(e0, e1, x) = t * t * (3 - 2 t) with
// t = clamp((x - e0) / (e1 - e0), 0, 1), exp is hdsr_expf and pow(x, y) = hdsr_expf(y * hdsr_logf(x)).
#include "hdsr_internal.h"

namespace hdsr {
namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 vscale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 vnormalize(V3 a) {
  const float l = sqrtf(vdot(a, a));
  return v3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float half_lo(uint32_t w) { return hdsr_f16_to_f32(w & 0xFFFFu); }
__device__ __forceinline__ float half_hi(uint32_t w) { return hdsr_f16_to_f32(w >> 16); }

// octDecode (gpu_screen_space_normals.wgsl:44-50, bilateral_normal_filter.wgsl:22-28)
__device__ __forceinline__ V3 oct_decode(float ex, float ey) {
  V3 n = v3(ex, ey, (1.0f - fabsf(ex)) - fabsf(ey));
  if (n.z < 0.0f) {
    const float sx = n.x >= 0.0f ? 1.0f : -1.0f, sy = n.y >= 0.0f ? 1.0f : -1.0f;
    const float ax = (1.0f - fabsf(n.y)) * sx, ay = (1.0f - fabsf(n.x)) * sy;
    n.x = ax;
    n.y = ay;
  }
  return vnormalize(n);
}

// octEncode + pack2x16float
__device__ __forceinline__ uint32_t oct_pack(V3 n) {
  const float sum = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
  float px = n.x / sum, py = n.y / sum;
  if (n.z < 0.0f) {
    const float sx = px >= 0.0f ? 1.0f : -1.0f, sy = py >= 0.0f ? 1.0f : -1.0f;
    const float nx = (1.0f - fabsf(py)) * sx, ny = (1.0f - fabsf(px)) * sy;
    px = nx;
    py = ny;
  }
  return hdsr_pack2x16(px, py);
}

// faceForwardNormal (gpu_screen_space_normals.wgsl:52-56)
__device__ __forceinline__ V3 face_forward(V3 normal, V3 view_dir) {
  const V3 n = vnormalize(normal), v = vnormalize(view_dir);
  return vdot(n, v) < 0.0f ? v3(-n.x, -n.y, -n.z) : n;
}

struct SsnArgs {
  const float* depth;
  const uint32_t* src;  // G-buffer normal words
  uint32_t* out;
  uint32_t w, h;
  float vw, vh;          // viewport
  float inv[16];         // inverse viewProj, column-major
  float cam[3];
  float mode, g0, g1;    // recoveryMode, grazingStart, grazingEnd
};

// reconstructViewPos (gpu_screen_space_normals.wgsl:26-34)
__device__ __forceinline__ V3 recon(const SsnArgs& a, float px, float py, float d) {
  const float nx = px / a.vw * 2.0f - 1.0f;
  const float ny = 1.0f - py / a.vh * 2.0f;
  const float* M = a.inv;
  const float wx = ((M[0] * nx + M[4] * ny) + M[8] * d) + M[12];
  const float wy = ((M[1] * nx + M[5] * ny) + M[9] * d) + M[13];
  const float wz = ((M[2] * nx + M[6] * ny) + M[10] * d) + M[14];
  const float ww = ((M[3] * nx + M[7] * ny) + M[11] * d) + M[15];
  return v3(wx / ww, wy / ww, wz / ww);
}

__global__ __launch_bounds__(256) void k_screen_space_normals(SsnArgs a) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= a.w || y >= a.h) return;
  const size_t o = (size_t)y * a.w + x;
  const float d = a.depth[o];
  if (d >= 0.9999f) {
    a.out[o] = a.mode > 0.5f ? a.src[o] : hdsr_pack2x16(0.0f, 0.0f);
    return;
  }
  const float cx = (float)x + 0.5f, cy = (float)y + 0.5f;
  const V3 pc = recon(a, cx, cy, d);
  const uint32_t xl = x ? x - 1 : 0, xr = x + 1 < a.w ? x + 1 : a.w - 1;
  const uint32_t yu = y ? y - 1 : 0, yd = y + 1 < a.h ? y + 1 : a.h - 1;
  const float dl = a.depth[(size_t)y * a.w + xl], dr = a.depth[(size_t)y * a.w + xr];
  const float du = a.depth[(size_t)yu * a.w + x], dd = a.depth[(size_t)yd * a.w + x];
  V3 ddx, ddy;
  if (fabsf(dl - d) < fabsf(dr - d)) ddx = vsub(pc, recon(a, cx - 1.0f, cy, dl));
  else ddx = vsub(recon(a, cx + 1.0f, cy, dr), pc);
  if (fabsf(du - d) < fabsf(dd - d)) ddy = vsub(pc, recon(a, cx, cy - 1.0f, du));
  else ddy = vsub(recon(a, cx, cy + 1.0f, dd), pc);
  const V3 view_dir = vnormalize(vsub(v3(a.cam[0], a.cam[1], a.cam[2]), pc));
  const V3 screen_n = face_forward(vnormalize(vcross(ddy, ddx)), view_dir);
  const uint32_t sw = a.src[o];
  const V3 source_n = face_forward(oct_decode(half_lo(sw), half_hi(sw)), view_dir);
  V3 n = screen_n;
  if (a.mode > 1.5f) {
    n = source_n;
  } else if (a.mode > 0.5f) {
    const float ndv = fmaxf(vdot(source_n, view_dir), 0.0f);
    const float t = fminf(fmaxf((ndv - a.g0) / (a.g1 - a.g0), 0.0f), 1.0f);
    const float blend = 1.0f - t * t * (3.0f - 2.0f * t);
    const float keep = 1.0f - blend;
    n = vnormalize(v3(source_n.x * keep + screen_n.x * blend, source_n.y * keep + screen_n.y * blend,
                      source_n.z * keep + screen_n.z * blend));
  }
  a.out[o] = oct_pack(n);
}

struct BilArgs {
  const uint32_t* in;
  const float* depth;
  uint32_t* out;
  uint32_t w, h;
  float spatial_sigma, depth_sigma, normal_sigma;
  int radius;
};

constexpr int kBilMaxR = 8;
constexpr int kBilHalo = 16 + 2 * kBilMaxR;

// One 16x16 block per workgroup.  The block's window (radius r on every side, coordinates clamped
// to the image as the reference's textureLoad is) is decoded once into LDS; each pixel then walks
// its (2r+1)^2 neighbours in the reference's dy-outer, dx-inner order.  The spatial weight depends
// only on (dx, dy) and is taken from a table built once per workgroup.
__global__ __launch_bounds__(256) void k_bilateral_normals(BilArgs a) {
  __shared__ float s_n[3][kBilHalo * kBilHalo];
  __shared__ float s_d[kBilHalo * kBilHalo];
  __shared__ float s_sw[(2 * kBilMaxR + 1) * (2 * kBilMaxR + 1)];
  const int r = a.radius, span = 16 + 2 * r, taps = 2 * r + 1;
  const int bx = (int)blockIdx.x * 16, by = (int)blockIdx.y * 16;
  for (int i = threadIdx.x; i < span * span; i += 256) {
    const int hx = i % span, hy = i / span;
    const int sx = min(max(bx + hx - r, 0), (int)a.w - 1), sy = min(max(by + hy - r, 0), (int)a.h - 1);
    const size_t o = (size_t)sy * a.w + sx;
    const uint32_t wd = a.in[o];
    const V3 n = oct_decode(half_lo(wd), half_hi(wd));
    s_n[0][i] = n.x;
    s_n[1][i] = n.y;
    s_n[2][i] = n.z;
    s_d[i] = a.depth[o];
  }
  for (int i = threadIdx.x; i < taps * taps; i += 256) {
    const int dx = i % taps - r, dy = i / taps - r;
    const float dist2 = (float)(dx * dx + dy * dy);
    s_sw[i] = hdsr_expf(-dist2 / (2.0f * a.spatial_sigma * a.spatial_sigma));
  }
  __syncthreads();
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int x = bx + lx, y = by + ly;
  if (x >= (int)a.w || y >= (int)a.h) return;
  const size_t o = (size_t)y * a.w + x;
  const uint32_t cw = a.in[o];
  const int ci = (ly + r) * span + (lx + r);
  const float cd = s_d[ci];
  if (cd >= 0.9999f) {
    a.out[o] = cw;
    return;
  }
  const V3 cn = v3(s_n[0][ci], s_n[1][ci], s_n[2][ci]);
  const float two_ds2 = 2.0f * a.depth_sigma * a.depth_sigma;
  const float inv_ns = 1.0f / fmaxf(a.normal_sigma, 0.01f);
  const float dn = fmaxf(cd, 0.001f);
  V3 acc = v3(0.0f, 0.0f, 0.0f);
  float tw = 0.0f;
  for (int dy = -r; dy <= r; ++dy) {
    for (int dx = -r; dx <= r; ++dx) {
      const int si = (ly + r + dy) * span + (lx + r + dx);
      const V3 sn = v3(s_n[0][si], s_n[1][si], s_n[2][si]);
      const float diff = fabsf(cd - s_d[si]) / dn;
      const float dw = hdsr_expf(-diff * diff / two_ds2);
      const float nd = fmaxf(vdot(cn, sn), 0.0f);
      const float nw = hdsr_expf(inv_ns * hdsr_logf(nd));
      const float wgt = (s_sw[(dy + r) * taps + (dx + r)] * dw) * nw;
      acc = v3(acc.x + sn.x * wgt, acc.y + sn.y * wgt, acc.z + sn.z * wgt);
      tw = tw + wgt;
    }
  }
  const float m = fmaxf(tw, 0.0001f);
  a.out[o] = oct_pack(vnormalize(v3(acc.x / m, acc.y / m, acc.z / m)));
}

}  // namespace

void launch_screen_space_normals(const float* depth, const uint32_t* src, uint32_t* out, uint32_t w, uint32_t h,
                                 const float* inv16, const float* cam3, uint32_t mode, float g0, float g1,
                                 hipStream_t stream) {
  if (!w || !h) return;
  SsnArgs a{};
  a.depth = depth;
  a.src = src;
  a.out = out;
  a.w = w;
  a.h = h;
  a.vw = (float)w;
  a.vh = (float)h;
  for (int i = 0; i < 16; ++i) a.inv[i] = inv16[i];
  for (int i = 0; i < 3; ++i) a.cam[i] = cam3[i];
  a.mode = (float)mode;
  a.g0 = g0;
  a.g1 = g1;
  hipLaunchKernelGGL(k_screen_space_normals, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, stream, a);
}

int launch_bilateral_normals(const uint32_t* in, const float* depth, uint32_t* out, uint32_t w, uint32_t h,
                             float spatial_sigma, float depth_sigma, float normal_sigma, int radius,
                             hipStream_t stream) {
  if (radius < 0 || radius > kBilMaxR) return -1;
  if (!w || !h) return 0;
  BilArgs a{in, depth, out, w, h, spatial_sigma, depth_sigma, normal_sigma, radius};
  hipLaunchKernelGGL(k_bilateral_normals, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, stream, a);
  return 0;
}

}  // namespace hdsr
