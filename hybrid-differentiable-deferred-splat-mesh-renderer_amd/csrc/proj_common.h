// proj_common.h - fp32 vector helpers and SH constants shared by the projection kernels
// (forward, and the backward of the differentiable mode).
//
// Arithmetic is IEEE fp32 with no contraction, in the operation order of the WGSL
// (src/shaders/gpu_project_splats.wgsl), so both kernels evaluate bit-identical forward values.
#pragma once
#include "hdsr_internal.h"

namespace hdsr {

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 smul(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ F3 vdiv(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float vlen(F3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ F3 normalize(F3 a) { return vdiv(a, vlen(a)); }

__device__ __forceinline__ void mat4_point(const float* m, F3 p, float out[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = ((m[i] * p.x + m[4 + i] * p.y) + m[8 + i] * p.z) + m[12 + i] * 1.0f;
}

// rotateAxis (gpu_project_splats.wgsl:75-79), q already divided by max(|q|, 1e-6)
__device__ __forceinline__ F3 rotate_axis(const float q[4], F3 axis) {
  const F3 u = f3(q[1], q[2], q[3]);
  const F3 inner = add(cross(u, axis), smul(q[0], axis));
  return add(axis, smul(2.0f, cross(u, inner)));
}

#define SH_C1 0.4886025119029199f
#define SH_C2_0 1.0925484305920792f
#define SH_C2_1 1.0925484305920792f
#define SH_C2_2 0.31539156525252005f
#define SH_C2_3 1.0925484305920792f
#define SH_C2_4 0.5462742152960396f
#define SH_C3_0 0.5900435899266435f
#define SH_C3_1 2.890611442640554f
#define SH_C3_2 0.4570457994644658f
#define SH_C3_3 0.3731763325901154f
#define SH_C3_4 0.4570457994644658f
#define SH_C3_5 1.445305721320277f
#define SH_C3_6 0.5900435899266435f

}  // namespace hdsr
