// k_composite.hip - shared-memory front-to-back rasterizer.
//
// Reference: composite (src/shaders/gpu_tile_splat_composite.wgsl:123-370).
//
// One 64-thread workgroup (= one wave64) per 16x16 tile; lane (lx, ly) = (lane & 7, lane >> 3)
// owns the 2x2 pixel quad at (16tx + 2lx, 16ty + 2ly).  Each batch stages 64 entries' records
// into LDS (one 32-byte record gather per lane); every lane then walks the batch in order.  The
// wave leaves the tile as soon as every pixel's transmittance is below the cutoff: from then on
// no splat can pass the 'T >= 0.001' gate, so the output is the one of the full walk.
//
// Per pixel and splat (fp32, no contraction, exp from hdsr_fmath.h):
//   power = (a*dx*dx + b*dx*dy) + c*dy*dy       a = -k0/2, b = -k1, c = -k2/2
//   alpha = min(0.99, opacity * exp(power))
//   valid = power > -4 && alpha > 1/255 && T >= 0.001
//   w = (alpha * T) * valid;  C += rgb * w;  T = valid ? T * (1 - alpha) : T
//   gw = w > 0.01 ? w : 0;    D += gw * depth;  W += gw
// Output: rgba = (C + T * (0.02, 0.02, 0.04), 1 - T), depth = D / max(W, 1e-4).
// With the G-buffer enabled, also the voted material / normal / confidence / emissive words.
#include "hdsr_internal.h"

namespace hdsr {
namespace {

__device__ __forceinline__ float h16(uint32_t bits) {
  _Float16 h;
  const uint16_t s = (uint16_t)(bits & 0xFFFFu);
  memcpy(&h, &s, 2);
  return (float)h;
}

__device__ __forceinline__ void oct_encode(float x, float y, float z, float* ox, float* oy) {
  const float sum = (fabsf(x) + fabsf(y)) + fabsf(z);
  float px = x / sum, py = y / sum;
  if (z < 0.0f) {
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

struct Px {
  float T, r, g, b, dw, ws;
};
struct Gb {
  float mr, mm, nx, ny, nz, cf, er, eg, eb;
};

template <bool kG>
__device__ __forceinline__ void blend(Px& p, Gb& q, float power, float o, float cr, float cg, float cb, float dep,
                                      const float4& ga, const float4& gb, float geb) {
  const float gauss = hdsr_expf(power);
  const float alpha = fminf(0.99f, o * gauss);
  const float newT = p.T * (1.0f - alpha);
  const bool valid = (power > -4.0f) && (alpha > (1.0f / 255.0f)) && (p.T >= 0.001f);
  const float w = (alpha * p.T) * (valid ? 1.0f : 0.0f);
  p.r = p.r + cr * w;
  p.g = p.g + cg * w;
  p.b = p.b + cb * w;
  p.T = valid ? newT : p.T;
  const float gw = w > 0.01f ? w : 0.0f;
  p.dw = p.dw + gw * dep;
  p.ws = p.ws + gw;
  if (kG) {
    q.mr = q.mr + ga.x * gw;
    q.mm = q.mm + ga.y * gw;
    q.nx = q.nx + gb.x * gw;
    q.ny = q.ny + gb.y * gw;
    q.nz = q.nz + gb.z * gw;
    q.cf = q.cf + ga.z * gw;
    q.er = q.er + ga.w * w;
    q.eg = q.eg + gb.w * w;
    q.eb = q.eb + geb * w;
  }
}

template <bool kG>
__device__ __forceinline__ void store_px(const FrameDesc& F, uint32_t X, uint32_t Y, const Px& p, const Gb& q) {
  if (X >= F.width || Y >= F.height) return;
  const size_t o = (size_t)Y * F.width + X;
  const float4 c = make_float4(p.r + p.T * 0.02f, p.g + p.T * 0.02f, p.b + p.T * 0.04f, 1.0f - p.T);
  reinterpret_cast<float4*>(F.rgba)[o] = c;
  const float safe = fmaxf(p.ws, 0.0001f);
  F.out_depth[o] = p.dw / safe;
  if (F.rgba16) {
    const uint2 h = make_uint2(hdsr_pack2x16(c.x, c.y), hdsr_pack2x16(c.z, c.w));
    reinterpret_cast<uint2*>(F.rgba16)[o] = h;
  }
  if (kG) {
    const float vx = q.nx, vy = q.ny + 0.0001f, vz = q.nz;
    const float l = sqrtf((vx * vx + vy * vy) + vz * vz);
    float ox, oy;
    oct_encode(vx / l, vy / l, vz / l, &ox, &oy);
    if (F.gb_normal) F.gb_normal[o] = hdsr_pack2x16(ox, oy);
    if (F.gb_material) {
      const uint4 m = make_uint4(hdsr_pack2x16(q.mr / safe, q.mm / safe), hdsr_pack2x16(q.er, q.eg),
                                 hdsr_pack2x16(q.eb, 0.0f), hdsr_pack2x16(q.cf / safe, 0.0f));
      reinterpret_cast<uint4*>(F.gb_material)[o] = m;
    }
  }
}

template <bool kG>
__global__ __launch_bounds__(64) void k_composite(Bufs b) {
  __shared__ float4 s_a[HDSR_BATCH];  // cx, cy, a, b
  __shared__ float4 s_b[HDSR_BATCH];  // c, opacity, r, g
  __shared__ float2 s_c[HDSR_BATCH];  // b, depth
  __shared__ float4 s_ga[kG ? HDSR_BATCH : 1];  // roughness, metalness, confidence, emissive r
  __shared__ float4 s_gb[kG ? HDSR_BATCH : 1];  // normal xyz, emissive g
  __shared__ float s_gc[kG ? HDSR_BATCH : 1];   // emissive b

  const uint32_t t = blockIdx.x;
  const uint32_t f = frame_of(b.tile_frame_start, b.num_frames, t);
  const FrameDesc& F = b.frames[f];
  const uint32_t lt = t - F.tile_base;
  const uint32_t tx = lt % F.tiles_x, ty = lt / F.tiles_x;
  const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
  const uint32_t start = b.tile_offsets[m];
  const uint32_t count = b.tile_counts[m];
  const uint32_t lane = threadIdx.x;
  const uint32_t X0 = tx * 16 + (lane & 7) * 2, Y0 = ty * 16 + (lane >> 3) * 2;
  const float p00x = (float)X0 + 0.5f, p00y = (float)Y0 + 0.5f;

  Px p[4];
  Gb q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    p[k] = Px{1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    q[k] = Gb{0, 0, 0, 0, 0, 0, 0, 0, 0};
  }

  for (uint32_t base = 0; base < count; base += HDSR_BATCH) {
    const uint32_t i = base + lane;
    if (i < count) {
      const uint32_t g = b.entries[start + i];
      const uint4 h0 = b.hot[(size_t)g * 2];
      const uint4 h1 = b.hot[(size_t)g * 2 + 1];
      s_a[lane] = make_float4(hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z) * -0.5f, -hdsr_u2f(h0.w));
      s_b[lane] = make_float4(hdsr_u2f(h1.x) * -0.5f, h16(h1.w >> 16), h16(h1.z), h16(h1.z >> 16));
      s_c[lane] = make_float2(h16(h1.w), hdsr_u2f(h1.y));
      if (kG) {
        const uint4 c0 = b.cold[(size_t)g * 2];
        const uint4 c1 = b.cold[(size_t)g * 2 + 1];
        // octDecode (gpu_tile_splat_composite.wgsl:143-149)
        const float ex = h16(c0.z), ey = h16(c0.z >> 16);
        float dnx = ex, dny = ey, dnz = (1.0f - fabsf(ex)) - fabsf(ey);
        if (dnz < 0.0f) {
          const float sx = dnx >= 0.0f ? 1.0f : -1.0f, sy = dny >= 0.0f ? 1.0f : -1.0f;
          const float ax = (1.0f - fabsf(dny)) * sx, ay = (1.0f - fabsf(dnx)) * sy;
          dnx = ax;
          dny = ay;
        }
        const float dl = sqrtf((dnx * dnx + dny * dny) + dnz * dnz);
        s_ga[lane] = make_float4(h16(c0.x), h16(c0.x >> 16), h16(c0.w), h16(c1.x));
        s_gb[lane] = make_float4(dnx / dl, dny / dl, dnz / dl, h16(c1.x >> 16));
        s_gc[lane] = h16(c1.y);
      }
    }
    __syncthreads();
    const uint32_t nb = min(HDSR_BATCH, count - base);
    for (uint32_t j = 0; j < nb; ++j) {
      const float4 A = s_a[j];
      const float4 B = s_b[j];
      const float2 C = s_c[j];
      const float dx = p00x - A.x, dy = p00y - A.y;
      const float dx1 = dx + 1.0f, dy1 = dy + 1.0f;
      const float pw0 = (A.z * dx * dx + A.w * dx * dy) + B.x * dy * dy;
      const float pw1 = (A.z * dx1 * dx1 + A.w * dx1 * dy) + B.x * dy * dy;
      const float pw2 = (A.z * dx * dx + A.w * dx * dy1) + B.x * dy1 * dy1;
      const float pw3 = (A.z * dx1 * dx1 + A.w * dx1 * dy1) + B.x * dy1 * dy1;
      if (pw0 <= -4.0f && pw1 <= -4.0f && pw2 <= -4.0f && pw3 <= -4.0f) continue;
      float4 ga = make_float4(0, 0, 0, 0), gb = make_float4(0, 0, 0, 0);
      float gc = 0.0f;
      if (kG) {
        ga = s_ga[j];
        gb = s_gb[j];
        gc = s_gc[j];
      }
      blend<kG>(p[0], q[0], pw0, B.y, B.z, B.w, C.x, C.y, ga, gb, gc);
      blend<kG>(p[1], q[1], pw1, B.y, B.z, B.w, C.x, C.y, ga, gb, gc);
      blend<kG>(p[2], q[2], pw2, B.y, B.z, B.w, C.x, C.y, ga, gb, gc);
      blend<kG>(p[3], q[3], pw3, B.y, B.z, B.w, C.x, C.y, ga, gb, gc);
    }
    __syncthreads();
    const bool done = p[0].T < 0.001f && p[1].T < 0.001f && p[2].T < 0.001f && p[3].T < 0.001f;
    if (__all(done)) break;
  }
  store_px<kG>(F, X0, Y0, p[0], q[0]);
  store_px<kG>(F, X0 + 1, Y0, p[1], q[1]);
  store_px<kG>(F, X0, Y0 + 1, p[2], q[2]);
  store_px<kG>(F, X0 + 1, Y0 + 1, p[3], q[3]);
}

}  // namespace

void launch_composite(const Bufs& b, hipStream_t stream, bool gbuffer) {
  if (!b.total_tiles) return;
  if (gbuffer)
    hipLaunchKernelGGL(k_composite<true>, dim3(b.total_tiles), dim3(64), 0, stream, b);
  else
    hipLaunchKernelGGL(k_composite<false>, dim3(b.total_tiles), dim3(64), 0, stream, b);
}

}  // namespace hdsr
