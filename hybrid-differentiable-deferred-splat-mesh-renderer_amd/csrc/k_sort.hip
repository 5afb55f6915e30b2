// k_sort.hip - per-tile depth order: bitonic sort on 20-bit log-depth keys, with the 128-bucket
// log-depth cascade for tiles holding more than 4096 entries.
//
// Reference: tile_depth_sort (src/shaders/gpu_tile_depth_sort.wgsl:28-79), bucket_sort
// (gpu_tile_bucket_sort.wgsl:41-144), chunk_sort / write_chunk_indirect
// (gpu_tile_chunk_sort.wgsl:36-97).
//
// Key of an entry at local index i (< 4096) of a small tile or of a chunk:
//   y   = (ln(max(depth, 1e-6)) - logMin) * (128 / logRange)       (the bucket coordinate)
//   q20 = min(u32(y * 2^13), 2^20 - 1),  key = (q20 << 12) | i
// with logMin / logRange from the min and max depth bits of the whole tile.  Keys are unique, so
// the bitonic network's result equals a stable sort by q20.  The bucket of an entry is
// min(u32(y), 127) = q20 >> 13.
//
// Large tiles: one 1024-thread workgroup finds the depth range, histograms the 128 buckets and
// writes the tile's entries in bucket order to scratch, stably (ranks inside a wave come from
// seven ballots, across waves from a per-bucket prefix).  The bucket-ordered list is then cut
// into consecutive 4096-entry chunks (the reference's greedy packing splits buckets exactly
// there) and each chunk is sorted like a small tile, reading scratch and writing the tile's
// range of the entry list.  The reference's thread-0 serial phases and atomic cursors are gone.
#include "hdsr_internal.h"

namespace hdsr {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

__global__ __launch_bounds__(1024) void k_bucket(Bufs b) {
  __shared__ uint32_t red_min[16], red_max[16];
  __shared__ uint32_t bbase[HDSR_NUM_BUCKETS], bcnt[HDSR_NUM_BUCKETS], cursor[HDSR_NUM_BUCKETS];
  __shared__ uint32_t wcnt[16][HDSR_NUM_BUCKETS];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t nl = b.stats->large_tiles;
  for (uint32_t l = blockIdx.x; l < nl; l += gridDim.x) {
    const LargeTile L = b.large_tiles[l];
    const uint32_t* src = b.entries + L.start;
    uint32_t dmn = 0xFFFFFFFFu, dmx = 0;
    for (uint32_t i = t; i < L.count; i += 1024) {
      const uint32_t d = b.depth[src[i]];
      dmn = min(dmn, d);
      dmx = max(dmx, d);
    }
    dmn = wave_min(dmn);
    dmx = wave_max(dmx);
    if (lane == 0) {
      red_min[w] = dmn;
      red_max[w] = dmx;
    }
    if (t < HDSR_NUM_BUCKETS) bcnt[t] = 0;
    __syncthreads();
    dmn = red_min[0];
    dmx = red_max[0];
#pragma unroll
    for (uint32_t q = 1; q < 16; ++q) {
      dmn = min(dmn, red_min[q]);
      dmx = max(dmx, red_max[q]);
    }
    float lmin, scale;
    hdsr_log_frame(dmn, dmx, &lmin, &scale);
    for (uint32_t i = t; i < L.count; i += 1024) {
      const uint32_t k = hdsr_bucket_of(hdsr_log_pos(b.depth[src[i]], lmin, scale));
      atomicAdd(&bcnt[k], 1u);
    }
    __syncthreads();
    if (t == 0) {
      uint32_t run = 0;
      for (uint32_t k = 0; k < HDSR_NUM_BUCKETS; ++k) {
        bbase[k] = run;
        run += bcnt[k];
      }
    }
    if (t < HDSR_NUM_BUCKETS) cursor[t] = 0;
    __syncthreads();
    for (uint32_t r0 = 0; r0 < L.count; r0 += 1024) {
      const uint32_t i = r0 + t;
      const bool valid = i < L.count;
      const uint32_t g = valid ? src[i] : 0u;
      const uint32_t k = valid ? hdsr_bucket_of(hdsr_log_pos(b.depth[g], lmin, scale)) : 0u;
      unsigned long long m = __ballot(valid);
#pragma unroll
      for (uint32_t bit = 0; bit < 7; ++bit) {
        const bool s = (k >> bit) & 1u;
        const unsigned long long bm = __ballot(s);
        m &= s ? bm : ~bm;
      }
      const uint32_t rank = __popcll(m & below);
      const bool last = valid && ((m >> lane) == 1ull);
      wcnt[t >> 7][t & 127] = 0;
      wcnt[8 + (t >> 7)][t & 127] = 0;
      __syncthreads();
      if (last) wcnt[w][k] = rank + 1;
      __syncthreads();
      if (t < HDSR_NUM_BUCKETS) {
        uint32_t run = cursor[t];
#pragma unroll
        for (uint32_t q = 0; q < 16; ++q) {
          const uint32_t c = wcnt[q][t];
          wcnt[q][t] = run;
          run += c;
        }
        cursor[t] = run;
      }
      __syncthreads();
      if (valid) b.scratch[L.scratch + bbase[k] + wcnt[w][k] + rank] = g;
      __syncthreads();
    }
    if (t == 0) {
      b.large_tiles[l].dmin = dmn;
      b.large_tiles[l].dmax = dmx;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t insert_zero_bit(uint32_t v, uint32_t bit) {
  const uint32_t mask = (1u << bit) - 1u;
  return ((v >> bit) << (bit + 1u)) | (v & mask);
}

__global__ __launch_bounds__(256) void k_segsort(Bufs b) {
  __shared__ uint32_t keys[HDSR_MAX_TILE_ENTRIES];
  __shared__ uint32_t vals[HDSR_MAX_TILE_ENTRIES];
  __shared__ uint32_t red_min[4], red_max[4];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const uint32_t nseg = b.stats->segments;
  for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const Seg S = b.segs[s];
    const uint32_t* src = (S.large == kNone ? b.entries : b.scratch) + S.src;
    const uint32_t count = S.count;
    uint32_t d[16];
    uint32_t dmn = 0xFFFFFFFFu, dmx = 0;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
      const uint32_t i = j * 256 + t;
      d[j] = 0;
      if (i < count) {
        const uint32_t g = src[i];
        vals[i] = g;
        d[j] = b.depth[g];
        dmn = min(dmn, d[j]);
        dmx = max(dmx, d[j]);
      }
    }
    float lmin, scale;
    if (S.large == kNone) {
      dmn = wave_min(dmn);
      dmx = wave_max(dmx);
      if (lane == 0) {
        red_min[w] = dmn;
        red_max[w] = dmx;
      }
      __syncthreads();
      dmn = min(min(red_min[0], red_min[1]), min(red_min[2], red_min[3]));
      dmx = max(max(red_max[0], red_max[1]), max(red_max[2], red_max[3]));
    } else {
      dmn = b.large_tiles[S.large].dmin;
      dmx = b.large_tiles[S.large].dmax;
    }
    hdsr_log_frame(dmn, dmx, &lmin, &scale);
    uint32_t P = 1;
    while (P < count) P <<= 1;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
      const uint32_t i = j * 256 + t;
      if (i < count) keys[i] = (hdsr_q20_of(hdsr_log_pos(d[j], lmin, scale)) << 12) | i;
      else if (i < P) keys[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        const uint32_t bit = __builtin_ctz(j);
        for (uint32_t c = t; c < (P >> 1); c += 256) {
          const uint32_t l = insert_zero_bit(c, bit);
          const uint32_t r = l | j;
          const uint32_t a = keys[l], bb = keys[r];
          const bool asc = (l & k) == 0;
          if ((a > bb) == asc) {
            keys[l] = bb;
            keys[r] = a;
          }
        }
        __syncthreads();
      }
    }
    uint32_t* dst = b.entries + S.dst;
    for (uint32_t i = t; i < count; i += 256) {
      const uint32_t key = keys[i];
      dst[i] = vals[key & 0xFFFu];
      if (b.keys20) b.keys20[S.dst + i] = key;
    }
    __syncthreads();
  }
}

}  // namespace

void launch_bucket(const Bufs& b, hipStream_t stream) {
  if (!b.total_tiles) return;
  const uint32_t grid = b.total_tiles < 1024u ? b.total_tiles : 1024u;
  hipLaunchKernelGGL(k_bucket, dim3(grid), dim3(1024), 0, stream, b);
}

void launch_segment_sort(const Bufs& b, hipStream_t stream) {
  if (!b.seg_cap) return;
  const uint32_t grid = b.seg_cap < 16384u ? b.seg_cap : 16384u;
  hipLaunchKernelGGL(k_segsort, dim3(grid), dim3(256), 0, stream, b);
}

}  // namespace hdsr
