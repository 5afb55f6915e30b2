// k_composite.hip - shared-memory front-to-back rasterizer.
//
// Reference: composite (src/shaders/gpu_tile_splat_composite.wgsl:123-370).
//
// One 256-thread workgroup (four wave64s) per 16x16 tile, one pixel per thread.  The reference
// gives each of 64 threads a 2x2 quad; on CDNA the tile's serial chain (every splat of a tile is
// walked in order by the same lanes) is what bounds the kernel once all tiles are resident, so
// spreading the 256 pixels over 256 lanes shortens the heaviest tile's chain four-fold.  Each
// batch stages 256 entries' 32-byte records into LDS (one gather per thread); each wave (an 8x8
// quadrant) walks only the entries whose conservative box reaches its quadrant, a wave whose 64
// pixels are all saturated skips the batch, and the workgroup stops once all four are.  From
// saturation on no splat can pass the 'T >= 0.001' gate, so the output equals the full walk.
// Pixels outside the viewport (the last tile row or column) count as saturated from the start.
//
// Per pixel and splat (fp32, no contraction, exp from hdsr_fmath.h), with the reference's quad
// arithmetic kept: d = p00 - centre, plus 1 for the odd column / row of the quad:
//   power = (a*dx*dx + b*dx*dy) + c*dy*dy       a = -k0/2, b = -k1, c = -k2/2
//   alpha = min(0.99, opacity * exp(power))
//   valid = power > -4 && alpha > 1/255 && T >= 0.001
//   w = (alpha * T) * valid;  C += rgb * w;  T = valid ? T * (1 - alpha) : T
//   gw = w > 0.01 ? w : 0;    D += gw * depth;  W += gw
// Output: rgba = (C + T * (0.02, 0.02, 0.04), 1 - T), depth = D / max(W, 1e-4).
// With the G-buffer enabled, also the voted material / normal / confidence / emissive words.
//
// f16 variant (src/shaders/gpu_tile_splat_composite_f16.wgsl:170-281): power, exp and alpha stay
// fp32; T, C and w are binary16 and every operation on them is one native f16 operation (single
// rounding, no contraction):
//   a16 = f16(alpha);  newT = T * (1 - a16);  w = (a16 * T) * valid;  C = C + rgb16 * w
//   valid = power > -4 && alpha > f32(1h/255h) && f32(T) >= f32(0.001h)
//   gw = f32(w) > 0.01 ? f32(w) : 0 (depth / G-buffer votes in fp32), emissive += e * f32(w)
// Output: rgb = f32(C + T * (0.02h, 0.02h, 0.04h)), alpha = 1 - f32(T).  The constants are the
// binary16 values nearest to the reference's literals: 1h/255h = 257/65536, 0.001h = 1049/2^20,
// 0.02h = 1311/65536, 0.04h = 1311/32768.
#include "hdsr_internal.h"
#include "raster_common.h"

#include <type_traits>

namespace hdsr {
namespace {

constexpr uint32_t kBatch = 256;

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
// f16 variant: transmittance and colour in binary16
struct Ph {
  _Float16 T, r, g, b;
  float dw, ws;
};
constexpr float kAlphaTh16 = 257.0f / 65536.0f;     // f16(1/255)
constexpr float kTCut16 = 1049.0f / 1048576.0f;     // f16(0.001)
constexpr float kBgRG16 = 1311.0f / 65536.0f;       // f16(0.02)
constexpr float kBgB16 = 1311.0f / 32768.0f;        // f16(0.04)

// Loads through global-address-space pointers (global_load, counted by vmcnt only; a flat load
// would also hold every lgkmcnt wait that precedes a barrier).
typedef unsigned int gu32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 gload4(const uint4* p) {
  const gu32x4 v = *(const __attribute__((address_space(1))) gu32x4*)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t gload(const uint32_t* p) {
  return *(const __attribute__((address_space(1))) uint32_t*)(p);
}

// The float4 record at byte offset off of a staged array.
__device__ __forceinline__ float4 rec16(const float4* a, uint32_t off) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a) + off);
}

__device__ __forceinline__ bool saturated(const Px& p) { return p.T < 0.001f; }
__device__ __forceinline__ bool saturated(const Ph& p) { return p.T < (_Float16)kTCut16; }

// A deferred tile's pixel state between the rounds (planar: field f of pixel i at s[f*256 + i]);
// binary16 values convert to fp32 and back exactly.
template <class P>
__device__ __forceinline__ void save_px(float* s, const P& p) {
  s[0] = (float)p.T;
  s[256] = (float)p.r;
  s[512] = (float)p.g;
  s[768] = (float)p.b;
  s[1024] = p.dw;
  s[1280] = p.ws;
}
template <class P>
__device__ __forceinline__ void load_px(const float* s, P& p) {
  p.T = s[0];
  p.r = s[256];
  p.g = s[512];
  p.b = s[768];
  p.dw = s[1024];
  p.ws = s[1280];
}
struct Gb {
  float mr, mm, nx, ny, nz, cf, er, eg, eb;
};

// One entry into one pixel, without branches: an entry that does not pass the gate leaves every
// accumulator unchanged (w = 0, T kept).  exp is only consumed where power > -4, capped at 88; there
// hdsr_expf_gauss is the oracle's exponential bit for bit.  The cap is a signed integer min of the
// bits (one instruction, where a float min also canonicalises its input): for every float that is
// not a NaN it is min(power, 88).  Below -4 the value is unused: whatever it is there (any bits, a
// NaN or an infinity for an extreme argument), alpha = min(0.99, o * gauss) stays finite (min drops
// a NaN), so w = alpha * T * 0 is zero and T is kept.
__device__ __forceinline__ float alpha_of(float power, float o) {
  const float x = hdsr_u2f((uint32_t)min((int32_t)hdsr_f2u(power), (int32_t)0x42B00000));  // 88.0f
  return fminf(0.99f, o * hdsr_expf_gauss(x));
}

template <bool kG>
__device__ __forceinline__ bool blend(Px& p, Gb& q, float power, float alpha, float cr, float cg, float cb,
                                      float dep, const float4& ga, const float4& gb, float geb) {
  const float newT = p.T * (1.0f - alpha);
  const bool valid = (power > -4.0f) && (alpha > (1.0f / 255.0f)) && (p.T >= 0.001f);
  // (alpha * T) * (valid ? 1 : 0) as a select: alpha * T is finite and >= +0 (alpha = min(0.99, .)
  // with a non-negative opacity), so both give the same bits
  const float w = valid ? alpha * p.T : 0.0f;
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
  return valid;
}

// The f16 variant's blend: the same gate, with T, C and w carried in binary16.
template <bool kG>
__device__ __forceinline__ bool blend(Ph& p, Gb& q, float power, float alpha, float cr, float cg, float cb,
                                      float dep, const float4& ga, const float4& gb, float geb) {
  const _Float16 a16 = (_Float16)alpha;
  const _Float16 newT = p.T * ((_Float16)1.0f - a16);
  const bool valid = (power > -4.0f) && (alpha > kAlphaTh16) && (p.T >= (_Float16)kTCut16);
  const _Float16 w = valid ? a16 * p.T : (_Float16)0.0f;
  p.r = p.r + (_Float16)cr * w;
  p.g = p.g + (_Float16)cg * w;
  p.b = p.b + (_Float16)cb * w;
  p.T = valid ? newT : p.T;
  const float w32 = (float)w;
  const float gw = w32 > 0.01f ? w32 : 0.0f;
  p.dw = p.dw + gw * dep;
  p.ws = p.ws + gw;
  if (kG) {
    q.mr = q.mr + ga.x * gw;
    q.mm = q.mm + ga.y * gw;
    q.nx = q.nx + gb.x * gw;
    q.ny = q.ny + gb.y * gw;
    q.nz = q.nz + gb.z * gw;
    q.cf = q.cf + ga.z * gw;
    q.er = q.er + ga.w * w32;
    q.eg = q.eg + gb.w * w32;
    q.eb = q.eb + geb * w32;
  }
  return valid;
}

// One staged entry (byte offset off, records A / B already read) into this lane's pixel.
template <bool kG, bool kD, typename P>
__device__ __forceinline__ void blend_slot(P& p, Gb& q, float pw, float alpha, const float4& B, uint32_t off,
                                           const float4* s_c, const float4* s_ga, const float4* s_gb,
                                           const float* s_gc, const uint32_t* s_idx, uint32_t* last) {
  const float4 C = rec16(s_c, off);
  const uint32_t sl = off >> 4;
  float4 ga = make_float4(0, 0, 0, 0), gb = make_float4(0, 0, 0, 0);
  float gc = 0.0f;
  if (kG) {
    ga = s_ga[sl];
    gb = s_gb[sl];
    gc = s_gc[sl];
  }
  const bool v = blend<kG>(p, q, pw, alpha, B.y, B.z, C.x, C.y, ga, gb, gc);
  if (kD && v) *last = s_idx[sl] + 1u;
}

__device__ __forceinline__ float4 out_rgba(const Px& p) {
  return make_float4(p.r + p.T * 0.02f, p.g + p.T * 0.02f, p.b + p.T * 0.04f, 1.0f - p.T);
}
__device__ __forceinline__ float4 out_rgba(const Ph& p) {
  const _Float16 bg = (_Float16)kBgRG16, bgb = (_Float16)kBgB16;
  return make_float4((float)(p.r + p.T * bg), (float)(p.g + p.T * bg), (float)(p.b + p.T * bgb),
                     1.0f - (float)p.T);
}

template <bool kG, typename P>
__device__ __forceinline__ void store_px(const FrameDesc& F, uint32_t X, uint32_t Y, const P& p, const Gb& q) {
  if (X >= F.width || Y >= F.height) return;
  const size_t o = (size_t)Y * F.width + X;
  const float4 c = out_rgba(p);
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

// kD: differentiable mode.  Records come from the fp32 table (drec) instead of the f16-packed
// one, and each pixel also keeps its final transmittance and one past the index (in its tile's
// list) of the last entry that blended, which bound the backward walk.
// kH: the f16 variant (binary16 transmittance and colour).
//
// Wave w owns the 8x8 quadrant (w & 1, w >> 1) of the tile.  While a batch is staged, every entry
// tests its ellipse (reach_quadrants: the box, then the smallest power-level over each quadrant,
// with rounding margins) against the four quadrants and is appended, in list order, to the index
// list of each quadrant it may reach; a wave then walks only its own list, so entries that cannot
// touch its pixels cost it nothing.
template <bool kG, bool kD>
struct TileLds {
  float4 s_a[kBatch];  // cx, cy, a, b
  float4 s_b[kBatch];  // opacity, r, g, c (the blend's three fields 16-byte aligned)
  float4 s_c[kBatch];  // b, depth (float4: the three records share one slot index)
  float4 s_ga[kG ? kBatch : 1];  // roughness, metalness, confidence, emissive r
  float4 s_gb[kG ? kBatch : 1];  // normal xyz, emissive g
  float s_gc[kG ? kBatch : 1];   // emissive b
  uint32_t s_idx[kD ? kBatch : 1];  // entry index of each compacted slot
  uint16_t s_list[4][kBatch];       // per quadrant: 16 x the slots that may reach it, in order
  uint32_t s_cnt[5][4];             // per staging wave: kept entries, then per quadrant
  int s_done[4];
};

// One tile (global index t, uniform) with the workgroup's shared arrays L; the records of kDepth
// batches ahead are in flight (the launches use 1: many tiles share a CU).
template <bool kG, bool kD, bool kH, uint32_t kDepth>
__device__ __forceinline__ void composite_tile(const Bufs& b, const uint32_t t, TileLds<kG, kD>& L) {
  float4* const s_a = L.s_a;
  float4* const s_b = L.s_b;
  float4* const s_c = L.s_c;
  float4* const s_ga = L.s_ga;
  float4* const s_gb = L.s_gb;
  float* const s_gc = L.s_gc;
  uint32_t* const s_idx = L.s_idx;
  uint16_t(*const s_list)[kBatch] = L.s_list;
  uint32_t(*const s_cnt)[4] = L.s_cnt;
  int* const s_done = L.s_done;
  const uint32_t f = frame_of(b.tile_frame_start, b.num_frames, t);
  const FrameDesc& F = b.frames[f];
  const uint32_t lt = t - F.tile_base;
  const uint32_t tx = lt % F.tiles_x, ty = lt / F.tiles_x;
  const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
  const uint32_t start = b.tile_offsets[m];
  const uint32_t full = b.tile_counts[m];
  const uint32_t tid = threadIdx.x;
  // Deferred chunk sort: round 1 stops a large tile at the end of its first chunk (the only one
  // sorted yet); k_composite_resume resumes the tiles left unsaturated there from their saved
  // state.  The batches are those of one pass (4096 is a multiple of kBatch), so the pixels are
  // the same.  (Never with the G-buffer or the differentiable records: those frames run in one
  // pass.)
  const bool deferrable = !kG && !kD && b.round == 1 && full > HDSR_MAX_TILE_ENTRIES;  // uniform
  const uint32_t count = deferrable ? HDSR_MAX_TILE_ENTRIES : full;
  const uint32_t base0 = 0;
  float* const pst = b.px_state + (size_t)t * (256 * 6) + tid;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t lx = (wave & 1u) * 8u + (lane & 7u), ly = (wave >> 1) * 8u + (lane >> 3);
  const uint32_t X = tx * 16 + lx, Y = ty * 16 + ly;
  // the reference's quad origin and the +1 offsets of the quad's odd column / row
  const float p00x = (float)(X & ~1u) + 0.5f, p00y = (float)(Y & ~1u) + 0.5f;
  // +1 on the odd column / row, +0 elsewhere: (p00 - c) + 0 equals p00 - c except that -0 turns
  // into +0, which leaves the power's value unchanged
  const float add_x = (lx & 1u) ? 1.0f : 0.0f, add_y = (ly & 1u) ? 1.0f : 0.0f;

  using PxT = typename std::conditional<kH, Ph, Px>::type;
  PxT p{};
  // a pixel outside the viewport is never stored: it starts saturated, so it never keeps the
  // tile's walk going (with T = 0 no entry passes the gate, and the stored pixels do not depend on it)
  p.T = (X < F.width && Y < F.height) ? 1.0f : 0.0f;
  uint32_t last = 0;
  Gb q = Gb{0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t loaded = 0;

  const float tile_x0 = (float)(tx * 16) + 0.5f, tile_y0 = (float)(ty * 16) + 0.5f;
  // Prefetch: while batch k is blended, the records of batches k+1 .. k+kDepth and the entry
  // indices of batches up to k+2*kDepth are in flight.  The batch loop is unrolled kDepth times:
  // slot s holds the batches k = s (mod kDepth), so no register with a load in flight is moved.
  // Past the end of the list a lane reads the list's last entry and gathers its record, which the
  // staging ignores: every load stays unpredicated, so none waits for a select.
  const uint4* const rec = kD ? reinterpret_cast<const uint4*>(b.drec) : b.hot;
  // round 1 reads a large tile's first chunk from first[] when k_first_chunks sorted it there; the
  // tile's list in entries[] then stays canonical until round 2 or 3 orders it
  const uint32_t* const ent = (b.round == 1 && deferrable && b.first_chunks)
                                  ? b.first + (size_t)b.tile_large[m] * HDSR_MAX_TILE_ENTRIES
                                  : b.entries + start;
  constexpr uint32_t kRec = kD ? 3 : 2;  // 16-byte words per record
  const uint4 z4 = make_uint4(0, 0, 0, 0);
  const uint32_t last_e = count - 1u;  // read only when count > base0 >= 0
  uint32_t gq[kDepth], ge[kDepth];  // entry index of slot s's batch, and of the batch kDepth later
  uint4 h0q[kDepth], h1q[kDepth], h2q[kDepth];
#pragma unroll
  for (uint32_t s = 0; s < kDepth; ++s) {
    gq[s] = ge[s] = 0u;
    h0q[s] = h1q[s] = h2q[s] = z4;
  }
  if (count > base0) {  // uniform: an empty walk reads nothing
#pragma unroll
    for (uint32_t s = 0; s < kDepth; ++s) {
      const uint32_t id = base0 + s * kBatch + tid;
      gq[s] = gload(ent + min(id, last_e));
      ge[s] = gload(ent + min(id + kDepth * kBatch, last_e));
    }
#pragma unroll
    for (uint32_t s = 0; s < kDepth; ++s) {
      h0q[s] = gload4(rec + (size_t)gq[s] * kRec);
      h1q[s] = gload4(rec + (size_t)gq[s] * kRec + 1);
      if (kD) h2q[s] = gload4(rec + (size_t)gq[s] * kRec + 2);
    }
  }
  bool finished = false;  // every pixel saturated
  uint32_t base = base0;
  while (base < count) {  // uniform
#pragma unroll
    for (uint32_t s = 0; s < kDepth; ++s) {
      if (base >= count) break;  // uniform
      const uint32_t i = base + tid;
      const uint4 h0 = h0q[s], h1 = h1q[s], h2 = h2q[s];
      const uint32_t g = gq[s];
      uint32_t qm = 0;  // quadrants the entry may reach
      if (i < count)
        qm = reach_quadrants(hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z), hdsr_u2f(h0.w), hdsr_u2f(h1.x), tile_x0,
                             tile_y0);
      // compact the kept entries into slots, in order, and list each under its quadrants
      const bool keep = qm != 0;
      const unsigned long long km = __ballot(keep);
      unsigned long long qb[4];
    #pragma unroll
      for (uint32_t k = 0; k < 4; ++k) qb[k] = __ballot((qm >> k) & 1u);
      if (lane == 0) {
        s_cnt[0][wave] = (uint32_t)__popcll(km);
    #pragma unroll
        for (uint32_t k = 0; k < 4; ++k) s_cnt[1 + k][wave] = (uint32_t)__popcll(qb[k]);
      }
      __syncthreads();
      if (keep) {
        uint32_t slot = (uint32_t)__popcll(km & below);
        for (uint32_t w2 = 0; w2 < wave; ++w2) slot += s_cnt[0][w2];
        s_a[slot] = make_float4(hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z) * -0.5f, -hdsr_u2f(h0.w));
        if (kD) {
          s_b[slot] = make_float4(hdsr_u2f(h1.y), hdsr_u2f(h1.z), hdsr_u2f(h1.w), hdsr_u2f(h1.x) * -0.5f);
          s_c[slot] = make_float4(hdsr_u2f(h2.x), hdsr_u2f(h2.y), 0.0f, 0.0f);
          s_idx[slot] = i;
        } else {
          s_b[slot] = make_float4(h16(h1.w >> 16), h16(h1.z), h16(h1.z >> 16), hdsr_u2f(h1.x) * -0.5f);
          s_c[slot] = make_float4(h16(h1.w), hdsr_u2f(h1.y), 0.0f, 0.0f);
        }
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
          s_ga[slot] = make_float4(h16(c0.x), h16(c0.x >> 16), h16(c0.w), h16(c1.x));
          s_gb[slot] = make_float4(dnx / dl, dny / dl, dnz / dl, h16(c1.x >> 16));
          s_gc[slot] = h16(c1.y);
        }
    #pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
          if ((qm >> k) & 1u) {
            uint32_t pos = (uint32_t)__popcll(qb[k] & below);
            for (uint32_t w2 = 0; w2 < wave; ++w2) pos += s_cnt[1 + k][w2];
            s_list[k][pos] = (uint16_t)(slot * 16u);  // byte offset of the slot's float4 records
          }
      }
      // uniform per wave: the walk below is a scalar loop
      const uint32_t nq = __builtin_amdgcn_readfirstlane(s_cnt[1 + wave][0] + s_cnt[1 + wave][1] +
                                                         s_cnt[1 + wave][2] + s_cnt[1 + wave][3]);
      __syncthreads();
      {  // slot s takes the batch kDepth later
        const uint32_t i2 = base + 2 * kDepth * kBatch + tid;
        gq[s] = ge[s];
        h0q[s] = gload4(rec + (size_t)ge[s] * kRec);
        h1q[s] = gload4(rec + (size_t)ge[s] * kRec + 1);
        if (kD) h2q[s] = gload4(rec + (size_t)ge[s] * kRec + 2);
        ge[s] = gload(ent + min(i2, last_e));
      }
      loaded += min(kBatch, count - base);
      if (!__all(saturated(p))) {
        // two entries per step: their exponentials are independent chains, blended in list order
        const uint16_t* lst = s_list[wave];
        for (uint32_t j = 0; j < nq; j += 2) {
          if ((j & 30u) == 30u && __all(saturated(p))) break;  // wave saturated
          const bool hb = j + 1u < nq;  // uniform
          const uint32_t oa = lst[j], ob = hb ? lst[j + 1u] : oa;
          const float4 Aa = rec16(s_a, oa), Ba = rec16(s_b, oa);
          const float4 Ab = rec16(s_a, ob), Bb = rec16(s_b, ob);
          const float dxa = (p00x - Aa.x) + add_x, dya = (p00y - Aa.y) + add_y;
          const float dxb = (p00x - Ab.x) + add_x, dyb = (p00y - Ab.y) + add_y;
          const float pwa = (Aa.z * dxa * dxa + Aa.w * dxa * dya) + Ba.w * dya * dya;
          const float pwb = (Ab.z * dxb * dxb + Ab.w * dxb * dyb) + Bb.w * dyb * dyb;
          const bool ra = __any(pwa > -4.0f), rb = hb && __any(pwb > -4.0f);  // uniform
          if (!ra && !rb) continue;  // no pixel of the wave is reached by either
          const float ala = alpha_of(pwa, Ba.x), alb = alpha_of(pwb, Bb.x);
          if (ra) blend_slot<kG, kD>(p, q, pwa, ala, Ba, oa, s_c, s_ga, s_gb, s_gc, s_idx, &last);
          if (rb) blend_slot<kG, kD>(p, q, pwb, alb, Bb, ob, s_c, s_ga, s_gb, s_gc, s_idx, &last);
        }
      }
      const int wave_done = __all(saturated(p));
      if (lane == 0) s_done[wave] = wave_done;
      __syncthreads();
      if (s_done[0] && s_done[1] && s_done[2] && s_done[3]) {
        finished = true;
        break;
      }
      base += kBatch;
    }
    if (finished) break;
  }
  if (b.round == 1 && deferrable && !finished) {  // uniform: resumed in round 2
    save_px(pst, p);
    if (tid == 0) {
      // list the tile and its chunks after the first (they follow the first chunk's segment)
      const uint32_t l = b.tile_large[m];
      const uint32_t k = atomicAdd(&b.stats->pend_tiles, 1u);
      b.pend_tiles[k] = t;
      b.pend_large[k] = l;
      b.pend_aux[2 * k] = 0u;  // round 2's per-tile count of loaded entries and finished quadrants
      b.pend_aux[2 * k + 1] = 0u;
      b.large_tiles[l].pending = 1u;
      const uint32_t seg0 = b.large_tiles[l].seg;
      const uint32_t later = (full - 1u) / HDSR_MAX_TILE_ENTRIES;
      const uint32_t at = atomicAdd(&b.stats->pend_segs, later);
      for (uint32_t q = 0; q < later; ++q) b.pend_segs[at + q] = seg0 + 1u + q;
      atomicAdd(&b.stats->composite_refs, loaded);
      atomicAdd(&b.stats->resumed_refs, full - HDSR_MAX_TILE_ENTRIES);  // sorted in round 2
      atomicMax(&b.stats->resumed_max, full);
    }
    return;
  }
  store_px<kG>(F, X, Y, p, q);
  if (kD && X < F.width && Y < F.height) {
    const size_t o = (size_t)Y * F.width + X;
    if (F.t_final) F.t_final[o] = (float)p.T;
    if (F.last_idx) F.last_idx[o] = last;
  }
  if (tid == 0 && b.stats) {
    atomicAdd(&b.stats->composite_refs, loaded);
    atomicMax(&b.stats->composite_max_refs, loaded);
  }
}

// Rounds 0 and 1: one workgroup per tile.
template <bool kG, bool kD, bool kH>
__global__ __launch_bounds__(256) void k_composite(Bufs b) {
  __shared__ TileLds<kG, kD> L;
  composite_tile<kG, kD, kH, 1>(b, blockIdx.x, L);
}

// Round 2 of a deferred frame (no G-buffer, no differentiable records): the tiles round 1 left
// unsaturated, resumed from their saved state.  There are few (tens per 1080p frame) and each is
// a chain of dependent batches, so a tile gets four workgroups of one wave, one per 8x8 quadrant:
// the quadrants run on different CUs, need no workgroup barrier, and each stops when its own 64
// pixels saturate.  k_resume_masks wrote each entry's reach_quadrants mask beside it (in qkey) up
// to list index HDSR_RESUME_MASKED, so there a quadrant reads a batch's 256 entries and masks
// (four per lane, coalesced) and gathers only the records of the entries that may reach it (a
// lane with nothing to gather reads the first record, one line for all of them); past it, the
// quadrant gathers every record and tests it (reach_quadrant).  Records are in flight kResumeDepth
// batches ahead, entries and masks twice that.  The kept entries are staged in order and walked
// with the arithmetic of round 1, so every pixel is the same.  The tile's count of loaded entries is the largest of its
// quadrants' (whole batches up to the batch where its last pixel saturates), as one workgroup
// counts it.
constexpr uint32_t kResumeDepth = 2;
template <bool kH>
__device__ __forceinline__ void resume_quadrant(const Bufs& b, const uint32_t k, const uint32_t t, const uint32_t qd,
                                                float4* s_a, float4* s_b, float4* s_c) {
  const uint32_t f = frame_of(b.tile_frame_start, b.num_frames, t);
  const FrameDesc& F = b.frames[f];
  const uint32_t lt = t - F.tile_base;
  const uint32_t tx = lt % F.tiles_x, ty = lt / F.tiles_x;
  const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
  const uint32_t start = b.tile_offsets[m];
  const uint32_t count = b.tile_counts[m];  // > 4096: round 1 stopped at 4096
  const uint32_t lane = threadIdx.x;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t lx = (qd & 1u) * 8u + (lane & 7u), ly = (qd >> 1) * 8u + (lane >> 3);
  const uint32_t X = tx * 16 + lx, Y = ty * 16 + ly;
  const float p00x = (float)(X & ~1u) + 0.5f, p00y = (float)(Y & ~1u) + 0.5f;
  const float add_x = (lx & 1u) ? 1.0f : 0.0f, add_y = (ly & 1u) ? 1.0f : 0.0f;
  using PxT = typename std::conditional<kH, Ph, Px>::type;
  PxT p{};
  load_px(b.px_state + (size_t)t * (256 * 6) + qd * 64 + lane, p);
  Gb gq = Gb{0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint4* const rec = b.hot;
  const uint32_t* const ent = b.entries + start;
  const uint32_t* const msk = b.qkey + start;
  const uint32_t last_e = count - 1u;
  constexpr uint32_t kD = kResumeDepth;
  // slot s: the kept entries (bit u: this lane's entry u) and their records, of the batches s
  // (mod kD); the entries and masks of the batch kD later
  uint32_t kq[kD], ga[kD][4], ma[kD][4];
  uint4 h0q[kD][4], h1q[kD][4];
  uint32_t base = HDSR_MAX_TILE_ENTRIES;
#pragma unroll
  for (uint32_t s = 0; s < kD; ++s) {
    uint32_t gg[4], mm[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t id = base + s * kBatch + u * 64 + lane;
      gg[u] = gload(ent + min(id, last_e));
      mm[u] = gload(msk + min(id, last_e));
      ga[s][u] = gload(ent + min(id + kD * kBatch, last_e));
      ma[s][u] = gload(msk + min(id + kD * kBatch, last_e));
    }
    const bool masked = base + s * kBatch < HDSR_RESUME_MASKED;  // uniform
    kq[s] = 0;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const bool keep = base + s * kBatch + u * 64 + lane < count && (!masked || ((mm[u] >> qd) & 1u));
      kq[s] |= keep ? 1u << u : 0u;
      const size_t r = keep ? (size_t)gg[u] * 2 : 0;
      h0q[s][u] = gload4(rec + r);
      h1q[s][u] = gload4(rec + r + 1);
    }
  }
  const float tile_x0 = (float)(tx * 16) + 0.5f, tile_y0 = (float)(ty * 16) + 0.5f;
  uint32_t loaded = 0;
  bool finished = false;
  while (base < count && !finished) {  // uniform
#pragma unroll
    for (uint32_t s = 0; s < kD; ++s) {
      if (base >= count) break;  // uniform
      // stage the kept entries in list order (past the masked range, the entries that pass the
      // quadrant's own test)
      const bool masked = base < HDSR_RESUME_MASKED;  // uniform
      uint32_t nk = 0;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint4 h0 = h0q[s][u], h1 = h1q[s][u];
        const bool keep = ((kq[s] >> u) & 1u) &&
                          (masked || reach_quadrant(qd, hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z),
                                                    hdsr_u2f(h0.w), hdsr_u2f(h1.x), tile_x0, tile_y0));
        const unsigned long long km = __ballot(keep);
        if (keep) {
          const uint32_t slot = nk + (uint32_t)__popcll(km & below);
          s_a[slot] = make_float4(hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z) * -0.5f, -hdsr_u2f(h0.w));
          s_b[slot] = make_float4(h16(h1.w >> 16), h16(h1.z), h16(h1.z >> 16), hdsr_u2f(h1.x) * -0.5f);
          s_c[slot] = make_float4(h16(h1.w), hdsr_u2f(h1.y), 0.0f, 0.0f);
        }
        nk += (uint32_t)__popcll(km);
      }
      // slot s takes the batch kD later: its kept entries' records, and the entries and masks of
      // the batch after that
      {
        const uint32_t b1 = base + kD * kBatch, b2 = base + 2 * kD * kBatch;
        const bool masked1 = b1 < HDSR_RESUME_MASKED;  // uniform
        kq[s] = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const bool keep = b1 + u * 64 + lane < count && (!masked1 || ((ma[s][u] >> qd) & 1u));
          kq[s] |= keep ? 1u << u : 0u;
          const size_t r = keep ? (size_t)ga[s][u] * 2 : 0;
          h0q[s][u] = gload4(rec + r);
          h1q[s][u] = gload4(rec + r + 1);
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          ga[s][u] = gload(ent + min(b2 + u * 64 + lane, last_e));
          ma[s][u] = gload(msk + min(b2 + u * 64 + lane, last_e));
        }
      }
      loaded += min(kBatch, count - base);
      __syncthreads();  // one wave: orders the staging before the walk
      for (uint32_t j = 0; j < nk; j += 2) {
        if ((j & 30u) == 30u && __all(saturated(p))) break;  // quadrant saturated
        const bool hb = j + 1u < nk;  // uniform
        const uint32_t oa = j * 16u, ob = hb ? oa + 16u : oa;
        const float4 Aa = rec16(s_a, oa), Ba = rec16(s_b, oa);
        const float4 Ab = rec16(s_a, ob), Bb = rec16(s_b, ob);
        const float dxa = (p00x - Aa.x) + add_x, dya = (p00y - Aa.y) + add_y;
        const float dxb = (p00x - Ab.x) + add_x, dyb = (p00y - Ab.y) + add_y;
        const float pwa = (Aa.z * dxa * dxa + Aa.w * dxa * dya) + Ba.w * dya * dya;
        const float pwb = (Ab.z * dxb * dxb + Ab.w * dxb * dyb) + Bb.w * dyb * dyb;
        const bool ra = __any(pwa > -4.0f), rb = hb && __any(pwb > -4.0f);  // uniform
        if (!ra && !rb) continue;
        const float ala = alpha_of(pwa, Ba.x), alb = alpha_of(pwb, Bb.x);
        if (ra) blend_slot<false, false>(p, gq, pwa, ala, Ba, oa, s_c, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rb) blend_slot<false, false>(p, gq, pwb, alb, Bb, ob, s_c, nullptr, nullptr, nullptr, nullptr, nullptr);
      }
      __syncthreads();  // the next batch restages the shared arrays
      if (__all(saturated(p))) {
        finished = true;
        break;
      }
      base += kBatch;
    }
  }
  store_px<false>(F, X, Y, p, gq);
  if (lane == 0 && b.stats) {
    // the tile's loaded entries: the largest of its quadrants', added by the last to finish
    atomicMax(&b.pend_aux[2 * k], loaded);
    __threadfence();
    if (atomicAdd(&b.pend_aux[2 * k + 1], 1u) == 3u) {
      __threadfence();
      const uint32_t most = atomicMax(&b.pend_aux[2 * k], 0u);
      atomicAdd(&b.stats->composite_refs, most);
      atomicMax(&b.stats->composite_max_refs, HDSR_MAX_TILE_ENTRIES + most);
    }
  }
}

template <bool kH>
__global__ __launch_bounds__(64) void k_composite_resume(Bufs b) {
  __shared__ float4 s_a[kBatch], s_b[kBatch], s_c[kBatch];
  const uint32_t n = 4u * b.stats->pend_tiles;
  for (uint32_t k = blockIdx.x; k < n; k += gridDim.x)
    resume_quadrant<kH>(b, k >> 2, __builtin_amdgcn_readfirstlane(b.pend_tiles[k >> 2]), k & 3u, s_a, s_b, s_c);
}

// Before the rounds run again over the same lists (encodeCompositeOnly): the round-2 lists and the
// rasterizer's counters start empty.
__global__ void k_reset_pending(DevStats* s) {
  s->pend_tiles = 0;
  s->pend_segs = 0;
  s->resumed_refs = 0;
  s->resumed_max = 0;
  s->composite_refs = 0;
  s->composite_max_refs = 0;
}

}  // namespace

void launch_reset_pending(const Bufs& b, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset_pending, dim3(1), dim3(1), 0, stream, b.stats);
}

void launch_composite(const Bufs& b, hipStream_t stream, bool gbuffer, bool f16) {
  if (!b.total_tiles) return;
  const dim3 block(256);
  if (b.round == 2) {  // deferred frames only: no G-buffer, no differentiable records
    // four one-wave workgroups per listed tile: a grid that holds them at once, without one per tile
    const uint32_t q = 4u * b.total_tiles;
    const dim3 grid(q > 4096u ? 4096u : q);
    if (f16)
      hipLaunchKernelGGL(k_composite_resume<true>, grid, dim3(64), 0, stream, b);
    else
      hipLaunchKernelGGL(k_composite_resume<false>, grid, dim3(64), 0, stream, b);
    return;
  }
  const dim3 grid(b.total_tiles);
  if (b.drec)
    hipLaunchKernelGGL((k_composite<false, true, false>), grid, block, 0, stream, b);
  else if (f16 && gbuffer)
    hipLaunchKernelGGL((k_composite<true, false, true>), grid, block, 0, stream, b);
  else if (f16)
    hipLaunchKernelGGL((k_composite<false, false, true>), grid, block, 0, stream, b);
  else if (gbuffer)
    hipLaunchKernelGGL((k_composite<true, false, false>), grid, block, 0, stream, b);
  else
    hipLaunchKernelGGL((k_composite<false, false, false>), grid, block, 0, stream, b);
}

}  // namespace hdsr
