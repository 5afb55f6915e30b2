// k_sort.hip - per-tile depth order on 20-bit log-depth keys, with the 128-bucket log-depth
// cascade for tiles holding more than 4096 entries.
//
// Reference: tile_depth_sort (src/shaders/gpu_tile_depth_sort.wgsl:28-79), bucket_sort
// (gpu_tile_bucket_sort.wgsl:41-144), chunk_sort / write_chunk_indirect
// (gpu_tile_chunk_sort.wgsl:36-97).
//
// Key of an entry at local index i (< 4096) of a small tile or of a chunk:
//   y   = (ln(max(depth, 1e-6)) - logMin) * (128 / logRange)       (the bucket coordinate)
//   q20 = min(u32(y * 2^13), 2^20 - 1),  key = (q20 << 12) | i
// with logMin / logRange from the min and max depth bits of the whole tile.  Keys are unique, so
// the reference's bitonic network orders them exactly as a stable sort on q20 does; that sort is
// done here as an LSD radix sort in LDS over q20 - min(q20 of the segment): a span of b bits takes
// ceil(b / 8) passes of near-equal digit widths (4 to 8 bits), so only the span's bits are ranked
// (a chunk of a large tile spans a fraction of its tile's range).  Equal digits are ranked inside a
// wave with one ballot per digit bit and a wave-private running count, so a pass needs five
// barriers of a four-wave workgroup instead of the bitonic network's 78 stages.
// The bucket of an entry is min(u32(y), 127) = q20 >> 13.
//
// Large tiles: one workgroup (512 or 1024 threads) writes the tile's entries in stable bucket
// order to scratch (ranks by ballots, per-bucket prefix across waves).  The bucket-ordered list is
// then cut into consecutive 4096-entry chunks (the reference's greedy packing splits buckets
// exactly there) and each chunk is sorted like a small tile, reading scratch and writing the
// tile's range of the entry list.
//
// Every entry's q20 arrives with it (qkey[]): the tile-list expansion evaluates it from the tile's
// depth range (min / max folded during the counting walk) and ln(max(depth, 1e-6)) (computed once
// per splat by the projection, bln[].y), so these kernels read one 4-byte key stream and do no log
// arithmetic.
#include "hdsr_internal.h"

namespace hdsr {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// The lanes of m whose kBits-bit digit equals this lane's: one ballot per digit bit, folded into
// the mask as m & ~(ballot ^ s), s = all ones where this lane's bit is set (one three-input bit
// operation per 32-bit half; truth table 0x90 with the operands in that order).
template <uint32_t kBits>
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, unsigned long long m) {
#pragma unroll
  for (uint32_t bit = 0; bit < kBits; ++bit) {
    const uint32_t s = (uint32_t)((int32_t)(d << (31u - bit)) >> 31);
    const unsigned long long bm = __ballot(s != 0u);
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)m, (uint32_t)bm, s, 0x90);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(m >> 32), (uint32_t)(bm >> 32), s, 0x90);
    m = ((unsigned long long)hi << 32) | lo;
  }
  return m;
}

// Running count of a group of lanes with equal digits (m: the group's lanes, this lane among
// them when valid): the group's highest lane adds the group's size to the wave-private counter
// and the value before the add reaches every lane of the group through a permute.  Rounds do not
// wait for one another: LDS atomics of one wave execute in issue order.
__device__ __forceinline__ uint32_t group_count(uint32_t* counter, unsigned long long m, bool valid, uint32_t lane) {
  const uint32_t leader = 63u - (uint32_t)__clzll(m | 1ull);
  uint32_t old = 0;
  if (valid && lane == leader) old = atomicAdd(counter, (uint32_t)__popcll(m));
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)old);
}

// Large tiles.  Wave w of the kBucketThreads-thread workgroup owns the contiguous slice
// [w*S, (w+1)*S) of the tile's entries.  Pass 1: per-wave bucket counts (bucket = q20 >> 13 of the entry's key;
// ranks inside a round from seven ballots, a wave-private running count across rounds).  Then one
// scan over (bucket, wave) gives every (wave, bucket) its first scratch slot, and pass 2 walks the
// slice in the same order and writes each entry (and its key) at slot + its running rank: a
// stable bucket order with two workgroup barriers per tile.
// kBucketThreads = 512 (eight waves per tile) gives each wave longer slices between the barriers
// and lets four workgroups share a CU: 0.245 / 0.205 / 0.237 ms per 1080p frame at 1024 / 512 /
// 256 threads.  Frames whose large tiles average more than 16k entries (a 4K view into a dense
// cluster: tiles of 100k entries) use 1024 threads, so the biggest tiles do not form the tail.
template <uint32_t kBucketThreads>
__global__ __launch_bounds__(kBucketThreads) void k_bucket(Bufs b) {
  constexpr uint32_t kBucketWaves = kBucketThreads / 64;
  __shared__ uint32_t wc[kBucketWaves][HDSR_NUM_BUCKETS];
  __shared__ uint32_t wsum[2];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t nl = b.stats->large_tiles;
  for (uint32_t l = blockIdx.x; l < nl; l += gridDim.x) {
    const LargeTile L = b.large_tiles[l];
    const uint32_t* src = b.entries + L.start;
    const uint32_t* qsrc = b.qkey + L.start;
    const uint32_t S = ((L.count + kBucketThreads - 1u) / kBucketThreads) * 64u;
    const uint32_t lo = w * S, hi = min(lo + S, L.count);
    for (uint32_t q = t; q < kBucketWaves * HDSR_NUM_BUCKETS; q += kBucketThreads) wc[q >> 7][q & 127] = 0;
    __syncthreads();
    // pass 1: wave-private bucket counts, four rounds of loads in flight
    for (uint32_t r0 = lo; r0 < hi; r0 += 256) {
      uint32_t kk[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t i = r0 + u * 64 + lane;
        kk[u] = i < hi ? qsrc[i] >> 13 : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const bool valid = r0 + u * 64 + lane < hi;
        const uint32_t k = kk[u];
        const unsigned long long m = match_digit<7>(k, __ballot(valid));
        if (valid && (m >> lane) == 1ull) wc[w][k] += (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    // scan over (bucket, wave), bucket-major: threads 0..127 own one bucket each
    uint32_t tot = 0, inc = 0;
    if (t < HDSR_NUM_BUCKETS) {
#pragma unroll
      for (uint32_t q = 0; q < kBucketWaves; ++q) tot += wc[q][t];
      inc = tot;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t x = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += x;
      }
      if (lane == 63) wsum[w] = inc;
    }
    __syncthreads();
    if (t < HDSR_NUM_BUCKETS) {
      uint32_t run = inc - tot + (w == 1 ? wsum[0] : 0u);
#pragma unroll
      for (uint32_t q = 0; q < kBucketWaves; ++q) {
        const uint32_t c = wc[q][t];
        wc[q][t] = run;
        run += c;
      }
    }
    __syncthreads();
    // pass 2: stable scatter in bucket order (entries and their keys)
    for (uint32_t r0 = lo; r0 < hi; r0 += 256) {
      uint32_t gg[4], qq[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t i = r0 + u * 64 + lane;
        gg[u] = i < hi ? src[i] : 0u;
        qq[u] = i < hi ? qsrc[i] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const bool valid = r0 + u * 64 + lane < hi;
        const uint32_t k = qq[u] >> 13;
        const unsigned long long m = match_digit<7>(k, __ballot(valid));
        const uint32_t rank = __popcll(m & below);
        const uint32_t old = group_count(&wc[w][k], m, valid, lane);
        if (valid) {
          b.scratch[L.scratch + old + rank] = gg[u];
          b.qkey_s[L.scratch + old + rank] = qq[u];
        }
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}

// Threads per segment-sort workgroup (16 keys per thread).  Four waves per workgroup keep the
// passes' barriers cheap; one 16 KB key array (the passes sort in place) and the register budget
// let kSortMinWaves workgroups share a CU, so the global loads and the barrier waits of one
// segment overlap the passes of others.
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortMinWaves = 5;
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortPer = HDSR_MAX_TILE_ENTRIES / kSortThreads;

// One LSD pass over keys[0..n), in place: stable by the digit (key >> shift) & (2^bits - 1).
// Wave w owns positions [w*S, min((w+1)*S, n)) (S a multiple of 64, at most 64 * kSortPer) and
// walks them in rounds of 64; equal digits inside a round are ranked by ballots, across rounds by
// a wave-private count.  Every key is read into registers before the barriers that precede the
// scatter, so the scatter can overwrite the same array.  Every thread of the workgroup must call
// it (it contains barriers).
template <uint32_t kBits>
__device__ __forceinline__ void lds_radix_pass(uint32_t* ks, uint32_t n, uint32_t S, uint32_t shift,
                                               uint32_t (*wc)[256], uint32_t* wsum) {
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t mask = (1u << kBits) - 1u;
  for (uint32_t q = t; q < kSortWaves * 256; q += kSortThreads) wc[q >> 8][q & 255] = 0;
  __syncthreads();
  const uint32_t lo = w * S, hi = min(lo + S, n);
  const uint32_t rounds = S / 64;  // the same for every wave: a segment of n keys needs ceil(n/256)
  uint32_t keys[kSortPer], pos[kSortPer];
#pragma unroll
  for (uint32_t r = 0; r < kSortPer; ++r) {
    pos[r] = 0xFFFFFFFFu;
    if (r >= rounds) continue;  // uniform
    const uint32_t p = lo + r * 64 + lane;
    const bool valid = p < hi;
    const uint32_t key = valid ? ks[p] : 0u;
    const uint32_t d = (key >> shift) & mask;
    const unsigned long long m = match_digit<kBits>(d, __ballot(valid));
    const uint32_t rank = __popcll(m & below);
    const uint32_t old = group_count(&wc[w][d], m, valid, lane);
    keys[r] = key;
    pos[r] = valid ? old + rank : 0xFFFFFFFFu;
  }
  __syncthreads();
  // digit-major exclusive scan over (digit, wave): digit t's base, then each wave's offset
  uint32_t tot = 0, inc = 0;
  if (t < 256) {
#pragma unroll
    for (uint32_t q = 0; q < kSortWaves; ++q) tot += wc[q][t];
    inc = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(inc, d, 64);
      if (lane >= (uint32_t)d) inc += v;
    }
    if (lane == 63) wsum[w] = inc;
  }
  __syncthreads();
  if (t < 256) {
    uint32_t run = inc - tot;
    for (uint32_t q = 0; q < w; ++q) run += wsum[q];
#pragma unroll
    for (uint32_t q = 0; q < kSortWaves; ++q) {
      const uint32_t c = wc[q][t];
      wc[q][t] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kSortPer; ++r)
    if (r < rounds && pos[r] != 0xFFFFFFFFu) ks[wc[w][(keys[r] >> shift) & mask] + pos[r]] = keys[r];
  __syncthreads();
}

// One pass of a digit width known only at run time (uniform): 4 to 8 bits.
__device__ __forceinline__ void radix_pass_w(uint32_t w, uint32_t* ks, uint32_t n, uint32_t S, uint32_t shift,
                                             uint32_t (*wc)[256], uint32_t* wsum) {
  switch (w) {
    case 4: lds_radix_pass<4>(ks, n, S, shift, wc, wsum); break;
    case 5: lds_radix_pass<5>(ks, n, S, shift, wc, wsum); break;
    case 6: lds_radix_pass<6>(ks, n, S, shift, wc, wsum); break;
    case 7: lds_radix_pass<7>(ks, n, S, shift, wc, wsum); break;
    default: lds_radix_pass<8>(ks, n, S, shift, wc, wsum); break;
  }
}

// Small tiles and chunks: one workgroup per segment (grid-stride).  Keys are built from the
// streamed q20 keys; the entries wait in registers during the passes and are then staged in the
// key array, where each sorted key's low 12 bits index them (small tiles are sorted in place in
// the entry list, so the source range cannot be read after the first write).
__global__ __launch_bounds__(kSortThreads, kSortMinWaves) void k_segsort(Bufs b) {
  __shared__ uint32_t ka[HDSR_MAX_TILE_ENTRIES];
  __shared__ uint32_t wc[kSortWaves][256];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t red[2][kSortWaves];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  // round 2 walks the chunks the first round listed (the later chunks of the tiles it left
  // unsaturated); the other rounds walk every segment
  const bool listed = b.round == 2;  // uniform
  const uint32_t nseg = listed ? b.stats->pend_segs : b.stats->segments;
  for (uint32_t k = blockIdx.x; k < nseg; k += gridDim.x) {
    const Seg S = b.segs[listed ? b.pend_segs[k] : k];
    if (b.round == 1 || b.round == 3) {  // uniform: deferred chunk sort
      bool take;
      if (S.large == kNone) {
        take = b.round == 1;
      } else {  // round 1: the first chunk; round 3: every other chunk (completes the lists of a
                // deferred frame)
        const bool first = S.dst == b.large_tiles[S.large].start;
        take = b.round == 1 ? first : !first;
      }
      if (!take) continue;
    }
    const uint32_t* src = (S.large == kNone ? b.entries : b.scratch) + S.src;
    const uint32_t* qsrc = (S.large == kNone ? b.qkey : b.qkey_s) + S.src;
    if (S.count & kSegKeepOrder) continue;  // uniform: copied unsorted by k_segcopy
    const uint32_t count = S.count;
    uint32_t vv[kSortPer], qq[kSortPer];
    uint32_t qmin = 0xFFFFFFFFu, qmax = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSortPer; ++r) {
      const uint32_t i = r * kSortThreads + t;
      vv[r] = 0;
      qq[r] = 0;
      if (i < count) {
        vv[r] = src[i];
        qq[r] = qsrc[i];
        qmin = min(qmin, qq[r]);
        qmax = max(qmax, qq[r]);
      }
    }
    qmin = wave_min(qmin);
    qmax = wave_max(qmax);
    if (lane == 0) {
      red[0][w] = qmin;
      red[1][w] = qmax;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kSortWaves; ++q) {
      qmin = min(qmin, red[0][q]);
      qmax = max(qmax, red[1][q]);
    }
    qmin = __builtin_amdgcn_readfirstlane(qmin);  // uniform: scalar registers
    qmax = __builtin_amdgcn_readfirstlane(qmax);
    // keys relative to the segment's smallest q20: only the bits of qmax - qmin are ranked
#pragma unroll
    for (uint32_t r = 0; r < kSortPer; ++r) {
      const uint32_t i = r * kSortThreads + t;
      if (i < count) ka[i] = ((qq[r] - qmin) << 12) | i;
    }
    const uint32_t span = qmax - qmin;  // uniform; 0 for a segment of one entry or equal keys
    const uint32_t qbits = span ? 32u - (uint32_t)__clz(span) : 0u;
    const uint32_t Sw = ((count + kSortThreads - 1) / kSortThreads) * 64;
    // ceil(qbits / 8) passes of near-equal digit widths (4 to 8 bits) cover exactly the key span
    const uint32_t np = (qbits + 7u) / 8u;
    const uint32_t w0 = np ? max((qbits + np - 1u) / np, 4u) : 0u;
    uint32_t low = 0;
    for (uint32_t ps = 0; ps < np; ++ps) {  // uniform
      const uint32_t w = ps + 1u < np ? w0 : max(qbits - low, 4u);
      radix_pass_w(w, ka, count, Sw, 12u + low, wc, wsum);
      low += w;
    }
    if (np == 0) __syncthreads();  // the keys written above are read below
    // each sorted key's low 12 bits index the segment's entries: take this thread's sorted keys,
    // stage the entries (waiting in registers since the first load) in the same array, then
    // gather them
    uint32_t sk[kSortPer];
#pragma unroll
    for (uint32_t r = 0; r < kSortPer; ++r) {
      const uint32_t i = r * kSortThreads + t;
      sk[r] = i < count ? ka[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kSortPer; ++r) {
      const uint32_t i = r * kSortThreads + t;
      if (i < count) ka[i] = vv[r];
    }
    __syncthreads();
    uint32_t* dst = b.entries + S.dst;
#pragma unroll
    for (uint32_t r = 0; r < kSortPer; ++r) {
      const uint32_t i = r * kSortThreads + t;
      if (i < count) {
        dst[i] = ka[sk[r] & 0xFFFu];
        if (b.keys20) b.keys20[S.dst + i] = sk[r] + (qmin << 12);
      }
    }
    __syncthreads();
  }
}

// With the reference's caps: the chunks past the first 1,024 keep their bucket order; they are
// copied from the scratch to their place in the entry list, each key holding its position.
__global__ __launch_bounds__(256) void k_segcopy(Bufs b) {
  const uint32_t nseg = b.stats->segments;
  for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const Seg S = b.segs[s];
    if (!(S.count & kSegKeepOrder)) continue;  // uniform
    const uint32_t count = S.count & ~kSegKeepOrder;
    for (uint32_t i = threadIdx.x; i < count; i += 256) {
      b.entries[S.dst + i] = b.scratch[S.src + i];
      if (b.keys20) b.keys20[S.dst + i] = (b.qkey_s[S.src + i] << 12) | i;
    }
  }
}

}  // namespace

void launch_bucket(const Bufs& b, hipStream_t stream) {
  if (!b.total_tiles) return;
  if (b.bucket_wide) {
    const uint32_t grid = b.total_tiles < 1024u ? b.total_tiles : 1024u;
    hipLaunchKernelGGL(k_bucket<1024>, dim3(grid), dim3(1024), 0, stream, b);
  } else {
    const uint32_t grid = b.total_tiles < 4096u ? b.total_tiles : 4096u;
    hipLaunchKernelGGL(k_bucket<512>, dim3(grid), dim3(512), 0, stream, b);
  }
}

void launch_segment_sort(const Bufs& b, hipStream_t stream) {
  if (!b.seg_cap) return;
  // round 2 sorts the listed chunks only: a grid that holds them at once without dispatching
  // thousands of workgroups that find nothing to do
  const uint32_t cap = b.round == 2 ? 1024u : 2048u * 1024u / kSortThreads;
  const uint32_t grid = b.seg_cap < cap ? b.seg_cap : cap;
  hipLaunchKernelGGL(k_segsort, dim3(grid), dim3(kSortThreads), 0, stream, b);
  if (b.ref_caps) hipLaunchKernelGGL(k_segcopy, dim3(grid), dim3(256), 0, stream, b);
}

}  // namespace hdsr
