// k_lighting.hip - deferred lighting: GGX specular + Lambertian diffuse over the G-buffer, with an
// ambient term, Reinhard tone mapping, sRGB encoding and additive emissive.
//
// Reference: main of src/shaders/gpu_deferred_lighting.wgsl:130-270 (helpers :42-108), encoded by
// src/splatRenderer.ts:446-548 (the 176-byte Params block :505-523) at :1115-1139.  This build takes
// the shader's path without an environment map (the reference's 1 x 1 placeholder, :138 / :224):
// background pixels pass the colour through, the ambient term is ambientColor * albedo * ao.
//
// Inputs per pixel: the compositor's colour as its rgba16float output texture holds it (the fp32
// result rounded to binary16), the fp32 G-buffer depth, the oct-normal word (after the normal
// passes) and the four material words; the AO texture's r channel (or 1) and a bilinear sample of
// the half-resolution bloom texture (or 0).  Output: rgba16float bits, round to nearest even.
//
// Arithmetic is IEEE fp32 without contraction in the WGSL's operation order: dot(a, b) =
// (a.x b.x + a.y b.y) + a.z b.z, normalize(v) = v / sqrt(dot(v, v)), mix(a, b, t) = a (1 - t) + b t,
// pow(x, y) = hdsr_expf(y * hdsr_logf(x)), i32(x) truncates toward zero (NaN -> 0, saturating).
// The bloom sample uses exact fp32 bilinear weights at texel centres (u * size - 0.5), U wrapping
// (repeat) and V clamped to the edge, as the reference's sampler is configured (src/ibl.ts:90-96).
#include "hdsr_internal.h"

namespace hdsr {
namespace {

struct L3 {
  float x, y, z;
};
__device__ __forceinline__ L3 l3(float x, float y, float z) { return L3{x, y, z}; }
__device__ __forceinline__ L3 ladd(L3 a, L3 b) { return l3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ L3 lsub(L3 a, L3 b) { return l3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ L3 lmul(L3 a, L3 b) { return l3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ L3 lscale(L3 a, float s) { return l3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ L3 ldivs(L3 a, float s) { return l3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ L3 ldiv(L3 a, L3 b) { return l3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ float ldot(L3 a, L3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ L3 lnorm(L3 a) { return ldivs(a, sqrtf(ldot(a, a))); }
__device__ __forceinline__ float lpow(float x, float y) { return hdsr_expf(y * hdsr_logf(x)); }
__device__ __forceinline__ float lsat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float lhalf(float x) { return hdsr_f16_to_f32(hdsr_f16_bits(x)); }

// WGSL i32(f32) then clamp(., 0, 255): the LUT texel a value in [0, 1] selects
__device__ __forceinline__ int lut_index(float x) {
  const float s = x * 255.0f;
  int i = 0;
  if (s != s) i = 0;
  else if (s <= -2147483648.0f) i = INT32_MIN;
  else if (s >= 2147483647.0f) i = INT32_MAX;
  else i = (int)s;
  return min(max(i, 0), 255);
}

// octDecode (gpu_deferred_lighting.wgsl:55-61)
__device__ __forceinline__ L3 oct_decode_l(float ex, float ey) {
  L3 n = l3(ex, ey, (1.0f - fabsf(ex)) - fabsf(ey));
  if (n.z < 0.0f) {
    const float sx = n.x >= 0.0f ? 1.0f : -1.0f, sy = n.y >= 0.0f ? 1.0f : -1.0f;
    const float ax = (1.0f - fabsf(n.y)) * sx, ay = (1.0f - fabsf(n.x)) * sy;
    n.x = ax;
    n.y = ay;
  }
  return lnorm(n);
}

struct LitArgs {
  LightingBlock p;         // the 44-float Params block (src/splatRenderer.ts:505-523)
  const float* rgba;       // compositor output, W*H*4 fp32
  const float* depth;      // W*H fp32
  const uint32_t* normal;  // W*H oct-normal words
  const uint32_t* mat;     // W*H*4 material words
  const float* ao;         // W*H*4 or null
  const float* bloom;      // bw*bh*4 or null
  const float* lut;        // 3*256: roughness, metalness, albedo curves
  uint16_t* out;           // W*H*4 binary16
  uint32_t w, h, bw, bh;
};

// textureSampleLevel(bloomTex, envSampler, uv, 0).rgb
__device__ __forceinline__ L3 bloom_sample(const LitArgs& a, float u, float v) {
  const float tx = u * (float)a.bw - 0.5f, ty = v * (float)a.bh - 0.5f;
  const float fx0 = floorf(tx), fy0 = floorf(ty);
  const float fx = tx - fx0, fy = ty - fy0;
  const int ix = (int)fx0, iy = (int)fy0;
  const int bw = (int)a.bw, bh = (int)a.bh;
  const int x0 = ((ix % bw) + bw) % bw, x1 = (((ix + 1) % bw) + bw) % bw;
  const int y0 = min(max(iy, 0), bh - 1), y1 = min(max(iy + 1, 0), bh - 1);
  const float4 t00 = reinterpret_cast<const float4*>(a.bloom)[(size_t)y0 * bw + x0];
  const float4 t10 = reinterpret_cast<const float4*>(a.bloom)[(size_t)y0 * bw + x1];
  const float4 t01 = reinterpret_cast<const float4*>(a.bloom)[(size_t)y1 * bw + x0];
  const float4 t11 = reinterpret_cast<const float4*>(a.bloom)[(size_t)y1 * bw + x1];
  const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
  return l3(((t00.x * w00 + t10.x * w10) + t01.x * w01) + t11.x * w11,
            ((t00.y * w00 + t10.y * w10) + t01.y * w01) + t11.y * w11,
            ((t00.z * w00 + t10.z * w10) + t01.z * w01) + t11.z * w11);
}

__device__ __forceinline__ void store_half4(uint16_t* out, size_t o, float r, float g, float b, float al) {
  const uint32_t lo = hdsr_f16_bits(r) | (hdsr_f16_bits(g) << 16);
  const uint32_t hi = hdsr_f16_bits(b) | (hdsr_f16_bits(al) << 16);
  reinterpret_cast<uint2*>(out)[o] = make_uint2(lo, hi);
}

constexpr float kPi = 3.14159265359f;
constexpr float kInvGamma = 0.454545454545f;  // 1.0 / 2.2 folded at full precision, then rounded to fp32

__global__ __launch_bounds__(256) void k_deferred_lighting(LitArgs a) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= a.w || y >= a.h) return;
  const size_t o = (size_t)y * a.w + x;
  const float* P = a.p.f;
  const float4 cs = reinterpret_cast<const float4*>(a.rgba)[o];
  const L3 albedo_srgb = l3(lhalf(cs.x), lhalf(cs.y), lhalf(cs.z));
  const float splat_opacity = lhalf(cs.w);
  const float depth = a.depth[o];
  if (depth >= 0.9999f || depth <= 0.0001f) {  // background, no environment map
    store_half4(a.out, o, albedo_srgb.x, albedo_srgb.y, albedo_srgb.z, 0.0f);
    return;
  }
  // albedo through its curve, then linearised (:171-174)
  const float albedo_lum = ldot(albedo_srgb, l3(0.2126f, 0.7152f, 0.0722f));
  const float albedo_remapped = a.lut[512 + lut_index(albedo_lum)];
  const float albedo_scale = albedo_lum < 0.001f ? 1.0f : albedo_remapped / fmaxf(albedo_lum, 0.001f);
  const L3 as = lscale(albedo_srgb, albedo_scale);
  const L3 albedo = l3(lpow(fmaxf(as.x, 0.0f), 2.2f), lpow(fmaxf(as.y, 0.0f), 2.2f), lpow(fmaxf(as.z, 0.0f), 2.2f));

  const uint32_t nw = a.normal[o];
  const L3 n_raw = oct_decode_l(hdsr_f16_to_f32(nw & 0xFFFFu), hdsr_f16_to_f32(nw >> 16));
  // reconstructWorldPos (:69-77)
  const float cx = (float)x + 0.5f, cy = (float)y + 0.5f;
  const float ndx = cx / P[0] * 2.0f - 1.0f, ndy = 1.0f - cy / P[1] * 2.0f;
  const float* M = P + 4;
  const float wx = ((M[0] * ndx + M[4] * ndy) + M[8] * depth) + M[12];
  const float wy = ((M[1] * ndx + M[5] * ndy) + M[9] * depth) + M[13];
  const float wz = ((M[2] * ndx + M[6] * ndy) + M[10] * depth) + M[14];
  const float ww = ((M[3] * ndx + M[7] * ndy) + M[11] * depth) + M[15];
  const L3 world = l3(wx / ww, wy / ww, wz / ww);
  const L3 V = lnorm(lsub(l3(P[20], P[21], P[22]), world));
  const L3 nn = lnorm(n_raw), vn = lnorm(V);  // faceForwardNormal (:63-67)
  const L3 N = ldot(nn, vn) < 0.0f ? l3(-nn.x, -nn.y, -nn.z) : nn;
  const L3 L = lnorm(l3(-P[24], -P[25], -P[26]));
  const L3 H = lnorm(ladd(V, L));
  const float NdotL = fmaxf(ldot(N, L), 0.0f);
  const float NdotV = fmaxf(ldot(N, V), 0.001f);
  const float NdotH = fmaxf(ldot(N, H), 0.0f);
  const float HdotV = fmaxf(ldot(H, V), 0.0f);

  // material (:189-194)
  const uint4 m4 = reinterpret_cast<const uint4*>(a.mat)[o];
  const float mrough = hdsr_f16_to_f32(m4.x & 0xFFFFu), mmetal = hdsr_f16_to_f32(m4.x >> 16);
  const float roughness = fmaxf(a.lut[lut_index(mrough)], 0.04f);
  const float metallic = a.lut[256 + lut_index(mmetal)];
  const float om = 1.0f - metallic;
  const L3 F0 = l3(0.04f * om + albedo.x * metallic, 0.04f * om + albedo.y * metallic, 0.04f * om + albedo.z * metallic);

  // Cook-Torrance (:80-108, :196-207)
  const float ga = roughness * roughness;
  const float a2 = ga * ga;
  const float dd = (NdotH * NdotH) * (a2 - 1.0f) + 1.0f;
  const float D = a2 / ((kPi * dd) * dd);
  const float rk = roughness + 1.0f;
  const float k = (rk * rk) / 8.0f;
  const float G = (NdotV / (NdotV * (1.0f - k) + k)) * (NdotL / (NdotL * (1.0f - k) + k));
  const float fp = lpow(lsat(1.0f - HdotV), 5.0f);
  const L3 F = l3(F0.x + (1.0f - F0.x) * fp, F0.y + (1.0f - F0.y) * fp, F0.z + (1.0f - F0.z) * fp);
  const L3 numerator = lscale(F, D * G);
  const float denominator = (4.0f * NdotV) * NdotL + 0.0001f;
  const L3 specular = ldivs(numerator, denominator);
  const L3 kD = lscale(lsub(l3(1.0f, 1.0f, 1.0f), F), om);

  // ambient occlusion, lifted by emissive and bloom (:210-230)
  const float ao_raw = a.ao ? reinterpret_cast<const float4*>(a.ao)[o].x : 1.0f;
  const float em_r = hdsr_f16_to_f32(m4.y & 0xFFFFu), em_g = hdsr_f16_to_f32(m4.y >> 16);
  const float em_b = hdsr_f16_to_f32(m4.z & 0xFFFFu);
  const float em_mag = fmaxf(em_r, fmaxf(em_g, em_b));
  L3 bl = l3(0.0f, 0.0f, 0.0f);
  if (a.bloom) bl = bloom_sample(a, cx / P[0], cy / P[1]);
  const float bloom_mag = ldot(bl, l3(0.2126f, 0.7152f, 0.0722f));
  const float em_lift = lsat((em_mag * P[36]) * 0.5f);
  const float bloom_lift = lsat(bloom_mag * 2.0f);
  const float ao = fmaxf(ao_raw, fmaxf(em_lift, bloom_lift));

  const L3 light = l3(P[28], P[29], P[30]);
  L3 color;
  if (P[35] > 0.5f) {
    color = lscale(lscale(lmul(specular, light), P[31]), NdotL);
  } else {
    const L3 diffuse = ldivs(lmul(kD, albedo), kPi);
    const L3 Lo = lscale(lscale(lmul(ladd(diffuse, specular), light), P[31]), NdotL);
    const L3 ambient = lscale(lmul(l3(P[32], P[33], P[34]), albedo), ao);
    color = ladd(ambient, Lo);
  }
  // Reinhard, sRGB, emissive after tone mapping (:254-267)
  const L3 exposed = lscale(color, P[40]);
  const L3 tm = ldiv(exposed, ladd(exposed, l3(1.0f, 1.0f, 1.0f)));
  L3 mapped = l3(lpow(fmaxf(tm.x, 0.0f), kInvGamma), lpow(fmaxf(tm.y, 0.0f), kInvGamma),
                 lpow(fmaxf(tm.z, 0.0f), kInvGamma));
  if (em_mag > P[37]) mapped = ladd(mapped, lscale(l3(em_r, em_g, em_b), P[36]));
  store_half4(a.out, o, mapped.x, mapped.y, mapped.z, splat_opacity);
}

}  // namespace

void launch_deferred_lighting(const LightingBlock& p, const float* rgba, const float* depth, const uint32_t* normal,
                              const uint32_t* mat, const float* ao, const float* bloom, const float* lut,
                              uint16_t* out, uint32_t w, uint32_t h, hipStream_t stream) {
  if (!w || !h) return;
  LitArgs a{};
  a.p = p;
  a.rgba = rgba;
  a.depth = depth;
  a.normal = normal;
  a.mat = mat;
  a.ao = ao;
  a.bloom = bloom;
  a.lut = lut;
  a.out = out;
  a.w = w;
  a.h = h;
  a.bw = (w + 1) / 2;
  a.bh = (h + 1) / 2;
  hipLaunchKernelGGL(k_deferred_lighting, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, stream, a);
}

}  // namespace hdsr
