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

// Of the tile's four 8x8 quadrants (bit k: quadrant (k & 1, k >> 1), pixel centres from (x0, y0)
// + 8 * (k & 1, k >> 1), 8 wide), those whose pixels the splat can reach: power > -4 needs
// q = d^T K d < 8 at some pixel centre d, and the smallest q over a quadrant's rectangle of
// centres is 0 when the centre lies inside it and otherwise sits on one of its four edges, where
// q is a parabola in the free coordinate minimised at -K_xy x / K_yy (clamped to the edge).  The
// level 8 is raised by the kernel's own rounding bound (as reach_box) and by a margin for this
// estimate's, so an entry any pixel evaluation would accept is never dropped; an entry whose
// ellipse misses a quadrant that its bounding box meets (the box's corners) is dropped from that
// quadrant's list.  Returns 0 when the splat cannot reach the tile.
// Quadrant k of the tile (pixel centres from (x0, y0) + 8 * (k & 1, k >> 1), 8 wide), given the
// splat's box bx from reach_box over the whole tile, exact (a proper ellipse) and the raised
// level; see reach_quadrants.
__device__ __forceinline__ bool reach_one_quadrant(uint32_t k, float cx, float cy, float k0, float k1, float k2,
                                                   float x0, float y0, const float4& bx, bool exact, float level) {
  const float qx0 = x0 + (float)((k & 1u) * 8u), qy0 = y0 + (float)((k >> 1) * 8u);
  if (bx.y < qx0 || bx.x > qx0 + 7.0f || bx.w < qy0 || bx.z > qy0 + 7.0f) return false;
  if (!exact) return true;
  const float X0 = qx0 - cx, X1 = qx0 + 7.0f - cx, Y0 = qy0 - cy, Y1 = qy0 + 7.0f - cy;
  float qmin = 0.0f;
  if (!(X0 <= 0.0f && X1 >= 0.0f && Y0 <= 0.0f && Y1 >= 0.0f)) {
    const float ax = X0 > 0.0f ? X0 : X1;  // the two edges facing the centre
    const float ay = Y0 > 0.0f ? Y0 : Y1;
    const float dy = fminf(fmaxf(-k1 * ax / k2, Y0), Y1);
    const float dx = fminf(fmaxf(-k1 * ay / k0, X0), X1);
    const float qa = (k0 * ax * ax + 2.0f * k1 * ax * dy) + k2 * dy * dy;
    const float qb = (k0 * dx * dx + 2.0f * k1 * dx * ay) + k2 * ay * ay;
    // an edge facing the centre only along one axis (the centre inside the other's range)
    // still bounds: both are the parabola minima on the rectangle's border lines
    qmin = (X0 <= 0.0f && X1 >= 0.0f) ? qb : (Y0 <= 0.0f && Y1 >= 0.0f) ? qa : fminf(qa, qb);
  }
  return qmin < level;
}

// The raised level of the quadrant tests over the 16x16 tile from (x0, y0).
__device__ __forceinline__ float reach_level(float cx, float cy, float k0, float k1, float k2, float x0, float y0) {
  const float DX = fmaxf(fabsf(x0 - cx), fabsf(x0 + 15.0f - cx)), DY = fmaxf(fabsf(y0 - cy), fabsf(y0 + 15.0f - cy));
  const float sabs = k0 * DX * DX + 2.0f * fabsf(k1) * DX * DY + k2 * DY * DY;
  return 8.0f + (16.0f * 5.9604645e-08f * 4.0f) * sabs + 1e-3f;
}

__device__ __forceinline__ uint32_t reach_quadrants(float cx, float cy, float k0, float k1, float k2, float x0,
                                                    float y0) {
  const float4 bx = reach_box(cx, cy, k0, k1, k2, x0, x0 + 15.0f, y0, y0 + 15.0f);
  if (bx.x > bx.y) return 0u;
  uint32_t qm = 0;
  const bool exact = k0 > 0.0f && k2 > 0.0f && bx.x != -INFINITY;  // a proper ellipse
  const float level = reach_level(cx, cy, k0, k1, k2, x0, y0);
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k)
    if (reach_one_quadrant(k, cx, cy, k0, k1, k2, x0, y0, bx, exact, level)) qm |= 1u << k;
  return qm;
}

// reach_quadrants' bit k alone.
__device__ __forceinline__ bool reach_quadrant(uint32_t k, float cx, float cy, float k0, float k1, float k2, float x0,
                                               float y0) {
  const float4 bx = reach_box(cx, cy, k0, k1, k2, x0, x0 + 15.0f, y0, y0 + 15.0f);
  if (bx.x > bx.y) return false;
  const bool exact = k0 > 0.0f && k2 > 0.0f && bx.x != -INFINITY;
  return reach_one_quadrant(k, cx, cy, k0, k1, k2, x0, y0, bx, exact, reach_level(cx, cy, k0, k1, k2, x0, y0));
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
