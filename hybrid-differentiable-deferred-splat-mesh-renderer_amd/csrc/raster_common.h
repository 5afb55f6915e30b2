// raster_common.h - tile culling shared by the forward and backward rasterizers.
#pragma once
#include "hdsr_internal.h"

namespace hdsr {

// The y-interval outside which the splat's power is <= -4 at every pixel of a tile whose pixel
// centres lie in [x0, x1] x [y0, y1], or an empty interval (lo > hi) when the splat cannot
// reach the tile at all.  Exact arithmetic: power = -q/2 with q = d^T K d, K = [[k0,k1],[k1,k2]],
// so power > -4 needs q < 8, an ellipse with half extents sqrt(8 K^-1_xx), sqrt(8 K^-1_yy).
// The kernel's fp32 power differs from the exact one by at most ~4 ulp of the sum of the
// magnitudes of its three terms (bounded here over the tile), so the level 8 is raised by
// twice that bound: culling then never drops a splat that the fp32 evaluation would accept.
// reach_box: the same bound as a box (x lo, x hi, y lo, y hi) of the ellipse, tested against the
// tile; an empty box (lo > hi in x) when the splat cannot reach the tile.  The rounding margin is
// taken over the whole tile, so the box is also valid for any part of it.
__device__ __forceinline__ float4 reach_box(float cx, float cy, float k0, float k1, float k2, float x0, float x1,
                                            float y0, float y1) {
  const float p02 = k0 * k2, p11 = k1 * k1;
  const float det = (p02 - p11) - 4.0f * 5.9604645e-08f * (p02 + p11);
  if (!(det > 0.0f) || !(k0 > 0.0f) || !(k2 > 0.0f)) return make_float4(-INFINITY, INFINITY, -INFINITY, INFINITY);
  const float DX = fmaxf(fabsf(x0 - cx), fabsf(x1 - cx)), DY = fmaxf(fabsf(y0 - cy), fabsf(y1 - cy));
  const float sabs = 0.5f * (k0 * DX * DX + 2.0f * fabsf(k1) * DX * DY + k2 * DY * DY);
  const float level = 8.0f + 2.0f * (8.0f * 5.9604645e-08f * sabs * 2.0f + 1e-4f);
  const float hx = sqrtf(level * k2 / det) * 1.0001f + 1e-3f;
  const float hy = sqrtf(level * k0 / det) * 1.0001f + 1e-3f;
  if (cx + hx < x0 || cx - hx > x1 || cy + hy < y0 || cy - hy > y1) return make_float4(1.0f, -1.0f, 1.0f, -1.0f);
  return make_float4(cx - hx, cx + hx, cy - hy, cy + hy);
}

__device__ __forceinline__ float2 reach_y(float cx, float cy, float k0, float k1, float k2, float x0, float x1,
                                          float y0, float y1) {
  // lower bound of det(K): the two products and the difference each round once
  const float p02 = k0 * k2, p11 = k1 * k1;
  const float det = (p02 - p11) - 4.0f * 5.9604645e-08f * (p02 + p11);
  if (!(det > 0.0f) || !(k0 > 0.0f) || !(k2 > 0.0f)) return make_float2(-INFINITY, INFINITY);
  const float DX = fmaxf(fabsf(x0 - cx), fabsf(x1 - cx)), DY = fmaxf(fabsf(y0 - cy), fabsf(y1 - cy));
  const float sabs = 0.5f * (k0 * DX * DX + 2.0f * fabsf(k1) * DX * DY + k2 * DY * DY);
  const float level = 8.0f + 2.0f * (8.0f * 5.9604645e-08f * sabs * 2.0f + 1e-4f);
  const float hx = sqrtf(level * k2 / det) * 1.0001f + 1e-3f;
  const float hy = sqrtf(level * k0 / det) * 1.0001f + 1e-3f;
  if (cx + hx < x0 || cx - hx > x1 || cy + hy < y0 || cy - hy > y1) return make_float2(1.0f, -1.0f);
  return make_float2(cy - hy, cy + hy);
}

}  // namespace hdsr
