// k_depthsort.hip - the global back-to-front order (sortedIndices) on the GPU.
//
// Reference: createViewDepthSortInput + encodeGpuSortPrototype (src/gpuSortPrototype.ts:4-55,
// 120-140, 275-287), which the renderer's encodeSort runs (src/splatRenderer.ts:1018-1022) to
// produce the sortedIndices buffer the compositor reads.  The CPU equivalent is
// sortSplatIdsBackToFront (src/splatSort.ts:32-55).
//
// Key of splat i: the view-space depth z = ((V[2] x + V[6] y) + V[10] z) + V[14] evaluated in
// double (as the reference's JavaScript does, from f32 inputs), negated and rounded to f32; the
// array is padded to a power of two P with (-inf, 0xFFFFFFFF).  The reference sorts it with a
// bitonic network in descending mode whose compare-exchange puts the smaller index first on equal
// keys in both directions.  That tie rule is not the reverse of itself between the network's
// ascending and descending runs, so for exactly equal keys the output order depends on the
// network (it is not "ties to the lower index"); the same network is therefore run here, stage
// for stage: all stages with stride j < 4096 of a 4096-element block in LDS, longer strides two per
// global pass.  For P = 2^20 that is 9 LDS kernels and 20 global passes (the reference issues 210
// dispatches).
#include "hdsr_internal.h"

namespace hdsr {
namespace {

constexpr uint32_t kBitonicBlock = 4096;
constexpr uint32_t kBitonicThreads = 1024;  // 2 pairs per thread per stage of a 4096-element block

// should_swap of gpuSortPrototype.ts:20-26
__device__ __forceinline__ bool should_swap(float lk, uint32_t li, float rk, uint32_t ri, bool ascending) {
  if (ascending) return (lk > rk) || ((lk == rk) && (li > ri));
  return (lk < rk) || ((lk == rk) && (li > ri));
}

// Descending mode: a pair at (i, i | j) of stage (k, j) compares ascending when (i & k) != 0.
__device__ __forceinline__ void compare_exchange(uint2& a, uint2& b, uint32_t i, uint32_t k) {
  const bool asc = (i & k) != 0u;
  if (should_swap(hdsr_u2f(a.x), a.y, hdsr_u2f(b.x), b.y, asc)) {
    const uint2 t = a;
    a = b;
    b = t;
  }
}

__global__ void k_depth_keys(const float* __restrict__ pos, uint32_t n, uint32_t padded, double v2, double v6,
                             double v10, double v14, uint2* a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= padded) return;
  if (i >= n) {
    a[i] = make_uint2(hdsr_f2u(-INFINITY), 0xFFFFFFFFu);
    return;
  }
  const double z = ((v2 * (double)pos[3 * (size_t)i] + v6 * (double)pos[3 * (size_t)i + 1]) +
                    v10 * (double)pos[3 * (size_t)i + 2]) + v14;
  a[i] = make_uint2(hdsr_f2u((float)(-z)), i);
}

// One global stage (k, j): thread t owns the pair whose lower position has bit j clear.
__global__ void k_bitonic_global(uint2* a, uint32_t pairs, uint32_t k, uint32_t j) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pairs) return;
  const uint32_t i = (t / j) * 2u * j + (t % j);
  uint2 l = a[i], r = a[i + j];
  compare_exchange(l, r, i, k);
  a[i] = l;
  a[i + j] = r;
}

// Two global stages (k, j) then (k, j/2) in one pass: thread t owns the four positions
// i + {0, j/2, j, 3j/2}, i with bits j and j/2 clear; all four share bit k (j < k), so one
// direction serves both stages.
__global__ void k_bitonic_global2(uint2* a, uint32_t quads, uint32_t k, uint32_t j) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= quads) return;
  const uint32_t h = j / 2;
  const uint32_t i = (t / h) * 2u * j + (t % h);
  uint2 x0 = a[i], x1 = a[i + h], x2 = a[i + j], x3 = a[i + j + h];
  compare_exchange(x0, x2, i, k);
  compare_exchange(x1, x3, i, k);
  compare_exchange(x0, x1, i, k);
  compare_exchange(x2, x3, i, k);
  a[i] = x0;
  a[i + h] = x1;
  a[i + j] = x2;
  a[i + j + h] = x3;
}

// All stages (k, j) with k in [k0, k1] (powers of two) and j < block, starting at j = j0 for k0,
// on blocks of 'block' consecutive elements held in LDS.
__global__ __launch_bounds__(kBitonicThreads) void k_bitonic_local(uint2* a, uint32_t block, uint32_t k0, uint32_t k1,
                                                                   uint32_t j0) {
  __shared__ uint2 s[kBitonicBlock];
  const uint32_t base = blockIdx.x * block;
  for (uint32_t q = threadIdx.x; q < block; q += kBitonicThreads) s[q] = a[base + q];
  __syncthreads();
  const uint32_t half = block / 2;
  for (uint32_t k = k0; k <= k1; k <<= 1) {
    for (uint32_t j = (k == k0 ? j0 : k / 2); j >= 1; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < half; t += kBitonicThreads) {
        const uint32_t li = (t / j) * 2u * j + (t % j);
        uint2 l = s[li], r = s[li + j];
        compare_exchange(l, r, base + li, k);
        s[li] = l;
        s[li + j] = r;
      }
      __syncthreads();
    }
  }
  for (uint32_t q = threadIdx.x; q < block; q += kBitonicThreads) a[base + q] = s[q];
}

__global__ void k_take_indices(const uint2* a, uint32_t n, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i].y;
}

}  // namespace

uint32_t depth_sort_padded(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

void launch_depth_sort(const float* positions, uint32_t n, const float* view16, uint32_t* out, uint2* work,
                       hipStream_t stream) {
  if (!n) return;
  const uint32_t P = depth_sort_padded(n);
  hipLaunchKernelGGL(k_depth_keys, dim3((P + 255) / 256), dim3(256), 0, stream, positions, n, P, (double)view16[2],
                     (double)view16[6], (double)view16[10], (double)view16[14], work);
  if (P > 1) {
    const uint32_t block = P < kBitonicBlock ? P : kBitonicBlock;
    const uint32_t nb = P / block;
    // every stage with k <= block is local to a block
    hipLaunchKernelGGL(k_bitonic_local, dim3(nb), dim3(kBitonicThreads), 0, stream, work, block, 2u, block, 1u);
    for (uint32_t k = block * 2; k <= P; k <<= 1) {
      uint32_t j = k / 2;
      for (; j >= block * 2; j >>= 2)  // strides j and j/2 together
        hipLaunchKernelGGL(k_bitonic_global2, dim3((P / 4 + 255) / 256), dim3(256), 0, stream, work, P / 4, k, j);
      if (j >= block)
        hipLaunchKernelGGL(k_bitonic_global, dim3((P / 2 + 255) / 256), dim3(256), 0, stream, work, P / 2, k, j);
      hipLaunchKernelGGL(k_bitonic_local, dim3(nb), dim3(kBitonicThreads), 0, stream, work, block, k, k, block / 2);
    }
  }
  hipLaunchKernelGGL(k_take_indices, dim3((n + 255) / 256), dim3(256), 0, stream, (const uint2*)work, n, out);
}

}  // namespace hdsr
