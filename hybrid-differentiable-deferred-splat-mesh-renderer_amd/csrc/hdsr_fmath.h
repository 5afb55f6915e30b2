// hdsr_fmath.h - scalar primitives that must round identically on the CPU and on gfx950.
//
// The compositor's integer outputs (tile keys, bucket indices, per-tile order) depend on
// ln(depth), and its pixel thresholds (alpha > 1/255, T >= 0.001, power > -4) depend on
// exp(power).  A one-ulp difference between two exp implementations flips such a threshold
// somewhere in a 1080p frame and moves a pixel by ~4e-3, so the kernels and every host-side
// restatement evaluate exp/ln with the functions below.  They use only IEEE-754 correctly
// rounded operations (+, -, *, /, fma, rint, ldexp), so a build without floating-point
// contraction (-ffp-contract=off on both hipcc and g++) produces the same bits on both sides.
//
// Methods:
//   hdsr_expf: Cody-Waite range reduction x = k*ln2 + r (two-constant split of ln2,
//              W. J. Cody & W. Waite, "Software Manual for the Elementary Functions", 1980),
//              then the degree-7 Taylor polynomial of e^r on |r| <= ln2/2 (truncation error
//              < 3e-9 relative), then an exact scaling by 2^k.
//   hdsr_logf: x = 2^e * m with m in [sqrt(1/2), sqrt(2)); ln m = 2*atanh(s) with
//              s = (m-1)/(m+1), summed as the odd series 2s(1 + s^2/3 + ... + s^10/11)
//              (|s| <= 0.1716, truncation error < 1e-10 relative); ln x = e*ln2 + ln m.
//
// Also here: the WGSL-style saturating float->u32 conversion, IEEE binary16 packing with
// round-to-nearest-even, and the 2D Morton code used for tile ids
// (reference: src/shaders/gpu_tile_splat_composite.wgsl:33-47).
#pragma once
#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define HDSR_HD __host__ __device__ static inline
#else
#define HDSR_HD static inline
#endif

#define HDSR_LN2_HI 0.693115234375f          /* 0x3F317000: 12 trailing zero bits */
#define HDSR_LN2_LO 3.19461832987e-05f       /* ln2 - HDSR_LN2_HI rounded to f32 */
#define HDSR_LOG2E 1.44269502162933350f

HDSR_HD uint32_t hdsr_f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

HDSR_HD float hdsr_u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// e^x for x in [-87, 88] (no range checks): the common path of hdsr_expf, bit-identical to it there.
HDSR_HD float hdsr_expf_core(float x) {
  const float k = rintf(x * HDSR_LOG2E);
  float r = fmaf(k, -HDSR_LN2_HI, x);
  r = fmaf(k, -HDSR_LN2_LO, r);
  float p = 1.98412698e-4f;          // 1/5040
  p = fmaf(p, r, 1.38888889e-3f);    // 1/720
  p = fmaf(p, r, 8.33333333e-3f);    // 1/120
  p = fmaf(p, r, 4.16666667e-2f);    // 1/24
  p = fmaf(p, r, 1.66666667e-1f);    // 1/6
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// e^x for the compositor's Gaussian weight, x in [-86, 88] (the compositor uses it where its power
// exceeds -4): the same reduction and polynomial as hdsr_expf_core with two steps cheaper.  k comes
// from one fma that adds 1.5 * 2^23, whose sum rounds the exact x * log2(e) to an integer (k in the
// low mantissa bits), and 2^k is added straight to the polynomial's exponent field (p lies in
// [0.7, 1.42], so the result stays a normal float for k in [-125, 127]).  Within 2 ulp of e^x on its
// range; it can differ from hdsr_expf in the last bit where x * log2(e) lies within rounding of a
// half-integer.  Outside the range the bits are meaningless (the compositor never reads them).
HDSR_HD float hdsr_expf_gauss(float x) {
  const float kf = fmaf(x, HDSR_LOG2E, 12582912.0f);
  const float k = kf - 12582912.0f;
  float r = fmaf(k, -HDSR_LN2_HI, x);
  r = fmaf(k, -HDSR_LN2_LO, r);
  float p = 1.98412698e-4f;          // 1/5040
  p = fmaf(p, r, 1.38888889e-3f);    // 1/720
  p = fmaf(p, r, 8.33333333e-3f);    // 1/120
  p = fmaf(p, r, 4.16666667e-2f);    // 1/24
  p = fmaf(p, r, 1.66666667e-1f);    // 1/6
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return hdsr_u2f(hdsr_f2u(p) + (hdsr_f2u(kf) << 23));
}

// e^x with <= 2 ulp error on [-87, 88]; 0 below -87, +inf above 88, NaN passes through.
HDSR_HD float hdsr_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) return INFINITY;
  if (x < -87.0f) return 0.0f;
  return hdsr_expf_core(x);
}

// ln x with <= 2 ulp error for finite x > 0; -inf at 0, NaN below 0, +inf at +inf.
HDSR_HD float hdsr_logf(float x) {
  if (x != x) return x;
  if (x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return x;
  uint32_t b = hdsr_f2u(x);
  int e = 0;
  if (b < 0x00800000u) {  // subnormal: scale by 2^23 (exact)
    x = x * 8388608.0f;
    b = hdsr_f2u(x);
    e = -23;
  }
  e += (int)(b >> 23) - 127;
  float m = hdsr_u2f((b & 0x007FFFFFu) | 0x3F800000u);  // [1, 2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;  // exact (Sterbenz)
  const float s = f / (2.0f + f);
  const float z = s * s;
  float p = 9.09090909e-2f;          // 1/11
  p = fmaf(p, z, 1.11111111e-1f);    // 1/9
  p = fmaf(p, z, 1.42857143e-1f);    // 1/7
  p = fmaf(p, z, 2.0e-1f);           // 1/5
  p = fmaf(p, z, 3.33333333e-1f);    // 1/3
  const float s2 = 2.0f * s;
  const float lnm = fmaf(s2 * z, p, s2);
  const float ef = (float)e;
  return fmaf(ef, HDSR_LN2_HI, fmaf(ef, HDSR_LN2_LO, lnm));
}

// WGSL u32(f32): truncate toward zero, saturate to [0, 2^32-1]; NaN -> 0.
HDSR_HD uint32_t hdsr_u32sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// IEEE binary16 from binary32, round to nearest even, subnormals kept, overflow to inf.
HDSR_HD uint32_t hdsr_f16_bits(float f) {
  const uint32_t x = hdsr_f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax > 0x7F800000u) return sign | 0x7E00u;        // NaN
  if (ax >= 0x477FF000u) return sign | 0x7C00u;       // >= 65520 rounds to inf
  if (ax >= 0x38800000u) {                            // normal half range
    const uint32_t mant = ax & 0x7FFFFFu;
    uint32_t h = ((((ax >> 23) - 127u + 15u)) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return sign | h;
  }
  if (ax < 0x33000000u) return sign;                  // < 2^-25 rounds to zero
  const uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
  const int e = (int)(ax >> 23) - 127;                // [-25, -15]
  const int sh = -e - 1;                              // [14, 24]
  uint32_t h = mant >> sh;
  const uint32_t rem = mant & ((1u << sh) - 1u);
  const uint32_t half = 1u << (sh - 1);
  if (rem > half || (rem == half && (h & 1u))) h++;
  return sign | h;
}

// binary16 bits -> binary32 (exact).
HDSR_HD float hdsr_f16_to_f32(uint32_t h) {
  const uint32_t sign = (h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1Fu;
  const uint32_t m = h & 0x3FFu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-08f;  // m * 2^-24, exact
    return hdsr_u2f(hdsr_f2u(v) | sign);
  }
  if (e == 31) return hdsr_u2f(sign | 0x7F800000u | (m << 13));
  return hdsr_u2f(sign | ((e - 15u + 127u) << 23) | (m << 13));
}

// WGSL pack2x16float.
HDSR_HD uint32_t hdsr_pack2x16(float a, float b) {
  return hdsr_f16_bits(a) | (hdsr_f16_bits(b) << 16);
}

// 16-bit x/y interleave, x in the even bits.
HDSR_HD uint32_t hdsr_morton2(uint32_t x, uint32_t y) {
  uint32_t mx = x & 0xFFFFu;
  mx = (mx | (mx << 8)) & 0x00FF00FFu;
  mx = (mx | (mx << 4)) & 0x0F0F0F0Fu;
  mx = (mx | (mx << 2)) & 0x33333333u;
  mx = (mx | (mx << 1)) & 0x55555555u;
  uint32_t my = y & 0xFFFFu;
  my = (my | (my << 8)) & 0x00FF00FFu;
  my = (my | (my << 4)) & 0x0F0F0F0Fu;
  my = (my | (my << 2)) & 0x33333333u;
  my = (my | (my << 1)) & 0x55555555u;
  return mx | (my << 1);
}

// Compositor constants (reference: gpu_project_splats.wgsl:21-25, gpu_tile_splat_composite.wgsl:23-28,130-132,
// gpu_tile_classify.wgsl:11, gpu_tile_bucket_sort.wgsl:13-14).
#define HDSR_PROJ_STRIDE 14u
#define HDSR_SENTINEL 0xFFFFFFFFu
#define HDSR_LARGE_SPLAT_TILES 16u
#define HDSR_MAX_TILE_ENTRIES 4096u
#define HDSR_NUM_BUCKETS 128u
#define HDSR_BATCH 64u
#define HDSR_TILE_PX 16u
#define HDSR_T_CUTOFF 0.001f
#define HDSR_ALPHA_MIN (1.0f / 255.0f)
#define HDSR_POWER_CUTOFF (-4.0f)
#define HDSR_ALPHA_MAX 0.99f
#define HDSR_LOW_PASS 0.3f
#define HDSR_SIGMA_RADIUS 3.0f
#define HDSR_FOOTPRINT_EPS 0.000000001f
#define HDSR_MIN_CLIP_W 0.0001f

// The tile's log-depth frame: logMin and the bucket scale 128/logRange (0 when the range is
// degenerate), from the min/max of the tile's u32 depth bits
// (reference: gpu_tile_bucket_sort.wgsl:75-79).
HDSR_HD void hdsr_log_frame(uint32_t dmin_bits, uint32_t dmax_bits, float* log_min, float* scale) {
  const float lmin = hdsr_logf(fmaxf(hdsr_u2f(dmin_bits), 1e-6f));
  const float lrange = hdsr_logf(fmaxf(hdsr_u2f(dmax_bits), 1e-6f)) - lmin;
  *log_min = lmin;
  *scale = (lrange < 1e-10f) ? 0.0f : 128.0f / lrange;
}

// Position of a depth in its tile's log frame, in bucket units (reference:
// gpu_tile_bucket_sort.wgsl:85).  bucket = min(u32(y), 127); the 20-bit key is
// q20 = min(u32(y * 2^13), 2^20 - 1), so q20 >> 13 == bucket for every depth.
HDSR_HD float hdsr_log_pos(uint32_t depth_bits, float log_min, float scale) {
  return (hdsr_logf(fmaxf(hdsr_u2f(depth_bits), 1e-6f)) - log_min) * scale;
}

HDSR_HD uint32_t hdsr_bucket_of(float y) {
  const uint32_t b = hdsr_u32sat(y);
  return b < HDSR_NUM_BUCKETS - 1u ? b : HDSR_NUM_BUCKETS - 1u;
}

// q20 of a position already multiplied by 2^13: min(u32sat(x), 2^20 - 1) without branches.  fmaxf
// sends NaN and negatives to 0, and the conversion of a value in [0, 2^20 - 1] truncates, as
// u32sat does.  Callers may fold the 2^13 into a finite tile scale (hdsr_log_frame's is 0 or
// at most 1.28e12): (d * scale) * 2^13 and d * (scale * 2^13) round alike (a power of two
// commutes with rounding), except where d * scale is subnormal or overflows, and there both
// give 0 or 2^20 - 1.
HDSR_HD uint32_t hdsr_q20_scaled(float x) {
  return (uint32_t)fminf(fmaxf(x, 0.0f), 1048575.0f);
}

HDSR_HD uint32_t hdsr_q20_of(float y) { return hdsr_q20_scaled(y * 8192.0f); }
