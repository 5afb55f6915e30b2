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
#include "raster_common.h"

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

// Adds the bucket counts (bucket = q20 >> 13) of keys [lo, hi) into one wave's 128 counters:
// per round of 64 keys the lanes of equal bucket are found by seven ballots and the group's highest
// lane adds the group's size, so no two lanes of a round update one counter (a plain LDS atomic per
// key serialises on the clustered depths of a tile).  The next 256 keys are in flight meanwhile.
__device__ __forceinline__ void wave_bucket_counts(const uint32_t* qsrc, uint32_t lo, uint32_t hi, uint32_t* cnt) {
  const uint32_t lane = __lane_id();
  uint32_t kk[4], kn[4];
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t i = lo + u * 64 + lane;
    kk[u] = i < hi ? qsrc[i] >> 13 : 0u;
  }
  for (uint32_t r0 = lo; r0 < hi; r0 += 256) {
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t i = r0 + 256 + u * 64 + lane;
      kn[u] = i < hi ? qsrc[i] >> 13 : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const bool valid = r0 + u * 64 + lane < hi;
      const uint32_t k = kk[u];
      const unsigned long long m = match_digit<7>(k, __ballot(valid));
      if (valid && (m >> lane) == 1ull) atomicAdd(&cnt[k], (uint32_t)__popcll(m));
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) kk[u] = kn[u];
  }
}

// Large tiles.  Wave w of the kBucketThreads-thread workgroup owns the contiguous slice
// [w*S, (w+1)*S) of the tile's entries.  Pass 1: per-wave bucket counts (bucket = q20 >> 13 of the
// entry's key; wave_bucket_counts).  Then one
// scan over (bucket, wave) gives every (wave, bucket) its first scratch slot, and pass 2 walks the
// slice in the same order and writes each entry (and its key) at slot + its running rank: a
// stable bucket order with two workgroup barriers per tile.
// kBucketThreads = 512 (eight waves per tile) gives each wave longer slices between the barriers
// and lets four workgroups share a CU: 0.245 / 0.205 / 0.237 ms per 1080p frame at 1024 / 512 /
// 256 threads.  Frames whose large tiles average more than 16k entries (a 4K view into a dense
// cluster: tiles of 100k entries) use 1024 threads, so the biggest tiles do not form the tail.
// Round 0 (one pass) orders every large tile; in a deferred frame, round 2 orders the tiles the
// first round left unsaturated (pend_large[]) and round 3 the others, whose canonical lists are
// still in entries[] (round 1 sorted their first chunks into first[] without a scatter).
template <uint32_t kBucketThreads>
__global__ __launch_bounds__(kBucketThreads) void k_bucket(Bufs b) {
  constexpr uint32_t kBucketWaves = kBucketThreads / 64;
  __shared__ uint32_t wc[kBucketWaves][HDSR_NUM_BUCKETS];
  __shared__ uint32_t wsum[2];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const bool listed = b.round == 2;  // uniform
  const uint32_t nl = listed ? b.stats->pend_tiles : b.stats->large_tiles;
  for (uint32_t k = blockIdx.x; k < nl; k += gridDim.x) {
    const uint32_t l = listed ? b.pend_large[k] : k;
    const LargeTile L = b.large_tiles[l];
    if (b.round == 3 && L.pending) continue;  // uniform: round 2 ordered it
    const uint32_t* src = b.entries + L.start;
    const uint32_t* qsrc = b.qkey + L.start;
    const uint32_t S = ((L.count + kBucketThreads - 1u) / kBucketThreads) * 64u;
    const uint32_t lo = w * S, hi = min(lo + S, L.count);
    for (uint32_t q = t; q < kBucketWaves * HDSR_NUM_BUCKETS; q += kBucketThreads) wc[q >> 7][q & 127] = 0;
    __syncthreads();
    // pass 1: wave-private bucket counts, the next 256 keys in flight
    wave_bucket_counts(qsrc, lo, hi, wc[w]);
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
// Wave w owns positions [w*S, min((w+1)*S, n)) (S a multiple of 64, at most 64 * 4096 / kT) and
// walks them in rounds of 64; equal digits inside a round are ranked by ballots, across rounds by
// a wave-private count.  Every key is read into registers before the barriers that precede the
// scatter, so the scatter can overwrite the same array.  Every thread of the workgroup must call
// it (it contains barriers).
// kT threads (kT / 64 waves, 4096 / kT keys each); wc has a row of kRow >= 2^bits digit counts per
// wave, and kRow <= kT.
template <uint32_t kBits, uint32_t kT, uint32_t kRow>
__device__ __forceinline__ void lds_radix_pass(uint32_t* ks, uint32_t n, uint32_t S, uint32_t shift,
                                               uint32_t (*wc)[kRow], uint32_t* wsum) {
  static_assert((1u << kBits) <= kRow && kRow <= kT && kRow % 64 == 0, "digit rows");
  constexpr uint32_t kWaves = kT / 64, kPer = HDSR_MAX_TILE_ENTRIES / kT;
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t mask = (1u << kBits) - 1u;
  for (uint32_t q = t; q < kWaves * kRow; q += kT) wc[q / kRow][q % kRow] = 0;
  __syncthreads();
  const uint32_t lo = w * S, hi = min(lo + S, n);
  const uint32_t rounds = S / 64;  // the same for every wave: a segment of n keys needs ceil(n/kT)
  uint32_t keys[kPer], pos[kPer];
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
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
  if (t < kRow) {  // the first kRow / 64 waves: one digit per thread
#pragma unroll
    for (uint32_t q = 0; q < kWaves; ++q) tot += wc[q][t];
    inc = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(inc, d, 64);
      if (lane >= (uint32_t)d) inc += v;
    }
    if (lane == 63) wsum[w] = inc;
  }
  __syncthreads();
  if (t < kRow) {
    uint32_t run = inc - tot;
    for (uint32_t q = 0; q < w; ++q) run += wsum[q];
#pragma unroll
    for (uint32_t q = 0; q < kWaves; ++q) {
      const uint32_t c = wc[q][t];
      wc[q][t] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r)
    if (r < rounds && pos[r] != 0xFFFFFFFFu) ks[wc[w][(keys[r] >> shift) & mask] + pos[r]] = keys[r];
  __syncthreads();
}

// One pass of a digit width known only at run time (uniform): 4 to log2(kRow) bits.
template <uint32_t kT, uint32_t kRow>
__device__ __forceinline__ void radix_pass_w(uint32_t w, uint32_t* ks, uint32_t n, uint32_t S, uint32_t shift,
                                             uint32_t (*wc)[kRow], uint32_t* wsum) {
  switch (w) {
    case 4: lds_radix_pass<4, kT, kRow>(ks, n, S, shift, wc, wsum); break;
    case 5: lds_radix_pass<5, kT, kRow>(ks, n, S, shift, wc, wsum); break;
    case 6: lds_radix_pass<6, kT, kRow>(ks, n, S, shift, wc, wsum); break;
    case 7: lds_radix_pass<7, kT, kRow>(ks, n, S, shift, wc, wsum); break;
    default:
      if constexpr (kRow >= 256) lds_radix_pass<8, kT, kRow>(ks, n, S, shift, wc, wsum);
      break;
  }
}

// Sorts the n (<= 4096) keys of ks in place in LDS by their bits 12 and up, which span qbits bits:
// ceil(qbits / B) passes of near-equal digit widths (4 to B bits, 2^B = kRow) cover exactly the
// key span.  Every thread of the kT-thread workgroup must call it; the sorted keys are visible on
// return.
template <uint32_t kT, uint32_t kRow>
__device__ __forceinline__ void lds_sort_span(uint32_t* ks, uint32_t n, uint32_t qbits, uint32_t (*wc)[kRow],
                                              uint32_t* wsum) {
  constexpr uint32_t kB = kRow >= 256 ? 8u : 7u;
  static_assert(kRow == 128 || kRow == 256, "digit rows");
  const uint32_t Sw = ((n + kT - 1) / kT) * 64;
  const uint32_t np = (qbits + kB - 1u) / kB;
  const uint32_t w0 = np ? max((qbits + np - 1u) / np, 4u) : 0u;
  uint32_t low = 0;
  for (uint32_t ps = 0; ps < np; ++ps) {  // uniform
    const uint32_t w = ps + 1u < np ? w0 : max(qbits - low, 4u);
    radix_pass_w<kT, kRow>(w, ks, n, Sw, 12u + low, wc, wsum);
    low += w;
  }
  if (np == 0) __syncthreads();  // the caller's key writes are read after the return
}

// One segment (a small tile or a chunk, count <= 4096 entries) sorted by its 20-bit keys with the
// kT threads of the workgroup: the entries and keys are read into registers, the keys taken
// relative to the segment's smallest q20 (only the bits of the span are ranked) and sorted in LDS
// (ka), then the entries are staged in the same array, where each sorted key's low 12 bits index
// them, and written to entries[dst, dst + count).  src may be that range (a small tile sorts in
// place: it cannot be read after the first write).  Every thread of the workgroup must call it;
// the shared arrays may be reused on return.
template <uint32_t kT, uint32_t kRow>
__device__ __forceinline__ void sort_segment(const Bufs& b, const uint32_t* src, const uint32_t* qsrc, uint32_t count,
                                             uint32_t dst_off, uint32_t* ka, uint32_t (*wc)[kRow], uint32_t* wsum,
                                             uint32_t (*red)[kT / 64]) {
  constexpr uint32_t kW = kT / 64, kPer = HDSR_MAX_TILE_ENTRIES / kT;
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  uint32_t vv[kPer], qq[kPer];
  uint32_t qmin = 0xFFFFFFFFu, qmax = 0;
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
    const uint32_t i = r * kT + t;
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
  for (uint32_t q = 0; q < kW; ++q) {
    qmin = min(qmin, red[0][q]);
    qmax = max(qmax, red[1][q]);
  }
  qmin = __builtin_amdgcn_readfirstlane(qmin);  // uniform: scalar registers
  qmax = __builtin_amdgcn_readfirstlane(qmax);
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
    const uint32_t i = r * kT + t;
    if (i < count) ka[i] = ((qq[r] - qmin) << 12) | i;
  }
  const uint32_t span = qmax - qmin;  // uniform; 0 for a segment of one entry or equal keys
  lds_sort_span<kT, kRow>(ka, count, span ? 32u - (uint32_t)__clz(span) : 0u, wc, wsum);
  uint32_t sk[kPer];
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
    const uint32_t i = r * kT + t;
    sk[r] = i < count ? ka[i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
    const uint32_t i = r * kT + t;
    if (i < count) ka[i] = vv[r];
  }
  __syncthreads();
  uint32_t* dst = b.entries + dst_off;
#pragma unroll
  for (uint32_t r = 0; r < kPer; ++r) {
    const uint32_t i = r * kT + t;
    if (i < count) {
      dst[i] = ka[sk[r] & 0xFFFu];
      if (b.keys20) b.keys20[dst_off + i] = sk[r] + (qmin << 12);
    }
  }
  __syncthreads();
}

// Small tiles and chunks: one workgroup per segment (grid-stride), kSortThreads threads.
// kT threads per segment: 256 where many segments are sorted (several workgroups share a CU),
// 1024 in round 2 after first chunks, whose few tens of chunks wait on the longest sort.
template <uint32_t kT>
__global__ __launch_bounds__(kT, kT == kSortThreads ? kSortMinWaves : 1) void k_segsort(Bufs b) {
  __shared__ uint32_t ka[HDSR_MAX_TILE_ENTRIES];
  __shared__ uint32_t wc[kT / 64][256];
  __shared__ uint32_t wsum[kT / 64];
  __shared__ uint32_t red[2][kT / 64];
  // round 2 walks the chunks the first round listed (the later chunks of the tiles it left
  // unsaturated); the other rounds walk every segment
  const bool listed = b.round == 2;  // uniform
  const uint32_t nseg = listed ? b.stats->pend_segs : b.stats->segments;
  for (uint32_t k = blockIdx.x; k < nseg; k += gridDim.x) {
    const Seg S = b.segs[listed ? b.pend_segs[k] : k];
    if (b.round == 1 || b.round == 3) {  // uniform: deferred chunk sort
      bool take;
      if (S.large == kNone) {
        take = b.round == 1 && !b.first_chunks;  // with first chunks, k_first_chunks sorts small tiles
      } else {
        const LargeTile& L = b.large_tiles[S.large];
        const bool head = S.dst == L.start;
        if (b.round == 1)  // the first chunk, unless k_first_chunks sorted it into first[]
          take = head && !b.first_chunks;
        else if (b.first_chunks)  // round 3 completes the lists: every chunk of a tile round 2 did
          take = !L.pending || head;  // not resume, the first chunk of one it did
        else  // round 3 after a scatter in round 1: every chunk after the first
          take = !head;
      }
      if (!take) continue;
    }
    const uint32_t* src = (S.large == kNone ? b.entries : b.scratch) + S.src;
    const uint32_t* qsrc = (S.large == kNone ? b.qkey : b.qkey_s) + S.src;
    if (S.count & kSegKeepOrder) continue;  // uniform: copied unsorted by k_segcopy
    sort_segment<kT, 256>(b, src, qsrc, S.count, S.dst, ka, wc, wsum, red);
  }
}

// The large tiles in order of decreasing size (by 4096-entry class), for the kernels that give a
// workgroup to each: with one tile per workgroup the dispatcher then schedules the longest first
// and the short ones fill in at the end.  One workgroup; the order inside a class is whatever the
// atomics give (it changes the schedule only, never a result).
__global__ __launch_bounds__(1024) void k_order_large(Bufs b) {
  __shared__ uint32_t hist[64];
  const uint32_t t = threadIdx.x, lane = __lane_id();
  const uint32_t nl = b.stats->large_tiles;
  if (t < 64) hist[t] = 0;
  __syncthreads();
  for (uint32_t l = t; l < nl; l += 1024) {
    const uint32_t c = b.large_tiles[l].count >> 12;
    atomicAdd(&hist[63u - min(c, 63u)], 1u);
  }
  __syncthreads();
  if (t < 64) {
    const uint32_t v = hist[t];
    hist[t] = wave_incl_scan(v) - v;
  }
  __syncthreads();
  for (uint32_t l = t; l < nl; l += 1024) {
    const uint32_t c = b.large_tiles[l].count >> 12;
    b.large_order[atomicAdd(&hist[63u - min(c, 63u)], 1u)] = l;
  }
  (void)lane;
}

// Round 1 of a deferred frame: the first chunk of large tile l, sorted, into first[l * 4096],
// without the bucket scatter.  The chunk holds bucket-order positions [0, 4096): every entry of
// the buckets below b*, the bucket holding position 4095, and the first 'need' entries of b* in
// list order (the reference's greedy packing splits b* there).  With the kT threads of the
// workgroup:
//   pass 1  per-wave bucket counts (wave_bucket_counts), then b*, 'need' and each wave's share;
//   pass 2  each wave walks its slice again and appends its chunk entries in list order (ranks by
//           two ballots per round), key (q20 << 12) | j at slot j, the entry at ea[j];
//   sort    the 4096 keys in LDS relative to the chunk's smallest q20 (as a segment, digits of at
//           most 7 bits, so the shared arrays leave room for four workgroups per CU).
// Within the chunk, equal q20 means one bucket, where the bucket order is the list order, so slot
// j orders ties as the bucket-order position does: the chunk is the one-pass chunk, entry for
// entry.  The tile's list in entries[] is left canonical for round 2 or 3.
// Both passes keep the next 256 keys of the slice in flight, and pass 2 stores a block's entries
// into LDS one block after gathering them.
constexpr uint32_t kFirstRow = HDSR_NUM_BUCKETS;  // 128 counts per wave: the buckets, then 7-bit digits
template <uint32_t kT>
struct FirstLds {
  uint32_t ka[HDSR_MAX_TILE_ENTRIES];
  uint32_t ea[HDSR_MAX_TILE_ENTRIES];
  uint32_t wc[kT / 64][kFirstRow];
  uint32_t wsum[kT / 64];
  uint32_t wsel[2][kT / 64];  // per wave: entries of the buckets below b*, entries of b*
  uint32_t cut[2];            // b*, need
  uint32_t red[2][kT / 64];
};

template <uint32_t kT>
__device__ __forceinline__ void first_chunk_tile(const Bufs& b, uint32_t l, FirstLds<kT>& S_) {
  constexpr uint32_t kW = kT / 64, kPer = HDSR_MAX_TILE_ENTRIES / kT, kN = HDSR_MAX_TILE_ENTRIES;
  constexpr uint32_t kRow = kFirstRow;
  uint32_t* const ka = S_.ka;
  uint32_t* const ea = S_.ea;
  uint32_t(*const wc)[kRow] = S_.wc;
  uint32_t* const wsum = S_.wsum;
  uint32_t(*const wsel)[kW] = S_.wsel;
  uint32_t* const cut = S_.cut;
  uint32_t(*const red)[kW] = S_.red;
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  {
    const LargeTile L = b.large_tiles[l];
    const uint32_t* src = b.entries + L.start;
    const uint32_t* qsrc = b.qkey + L.start;
    const uint32_t S = ((L.count + kT - 1u) / kT) * 64u;
    const uint32_t lo = min(w * S, L.count), hi = min(lo + S, L.count);
    for (uint32_t q = t; q < kW * kRow; q += kT) wc[q / kRow][q % kRow] = 0;
    if (t == 0) {
      cut[0] = HDSR_NUM_BUCKETS - 1u;
      cut[1] = 0u;
    }
    __syncthreads();
    // pass 1: wave-private bucket counts
    wave_bucket_counts(qsrc, lo, hi, wc[w]);
    __syncthreads();
    // b*: the bucket whose bucket-order range holds position 4095 (threads 0..127 own a bucket each)
    uint32_t tot = 0, inc = 0;
    if (t < HDSR_NUM_BUCKETS) {
#pragma unroll
      for (uint32_t q = 0; q < kW; ++q) tot += wc[q][t];
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
      const uint32_t ex = inc - tot + (w == 1 ? wsum[0] : 0u);
      if (tot != 0u && ex < kN && kN <= ex + tot) {
        cut[0] = t;
        cut[1] = kN - ex;
      }
    }
    __syncthreads();
    const uint32_t bs = __builtin_amdgcn_readfirstlane(cut[0]);
    const uint32_t need = __builtin_amdgcn_readfirstlane(cut[1]);
    {
      uint32_t below_bs = 0;
      for (uint32_t k = lane; k < HDSR_NUM_BUCKETS; k += 64)
        if (k < bs) below_bs += wc[w][k];
      below_bs = wave_sum(below_bs);
      if (lane == 0) {
        wsel[0][w] = below_bs;
        wsel[1][w] = wc[w][bs];
      }
    }
    __syncthreads();
    // this wave's first rank among the entries of b*, and its first slot in the chunk
    uint32_t runb = 0, runs = 0;
    for (uint32_t q = 0; q < w; ++q) {
      const uint32_t eq = wsel[1][q];
      runs += wsel[0][q] + min(eq, need > runb ? need - runb : 0u);
      runb += eq;
    }
    // pass 2: the chunk's entries in list order
    uint32_t qlo = 0xFFFFFFFFu, qhi = 0;
    {
      uint32_t qq[4], qn[4], pj[4], pg[4];
      bool ps[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t i = lo + u * 64 + lane;
        qq[u] = i < hi ? qsrc[i] : 0u;
        ps[u] = false;
        pj[u] = pg[u] = 0u;
      }
      for (uint32_t r0 = lo; r0 < hi; r0 += 256) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t i = r0 + 256 + u * 64 + lane;
          qn[u] = i < hi ? qsrc[i] : 0u;
        }
        uint32_t jj[4];
        bool sel[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const bool valid = r0 + u * 64 + lane < hi;
          const uint32_t k = qq[u] >> 13;
          const bool isb = valid && k == bs;
          const unsigned long long mb = __ballot(isb);
          const uint32_t rb = runb + (uint32_t)__popcll(mb & below);
          runb += (uint32_t)__popcll(mb);
          sel[u] = valid && (k < bs || (isb && rb < need));
          const unsigned long long ms = __ballot(sel[u]);
          jj[u] = runs + (uint32_t)__popcll(ms & below);
          runs += (uint32_t)__popcll(ms);
          sel[u] = sel[u] && jj[u] < kN;  // always: pass 1 counted the same keys
        }
        // the previous block's entries have had this block's loads and ranks to arrive
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
          if (ps[u]) ea[pj[u]] = pg[u];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          pg[u] = sel[u] ? src[r0 + u * 64 + lane] : 0u;
          pj[u] = jj[u];
          ps[u] = sel[u];
          if (sel[u]) {
            ka[jj[u]] = (qq[u] << 12) | jj[u];
            qlo = min(qlo, qq[u]);
            qhi = max(qhi, qq[u]);
          }
          qq[u] = qn[u];
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        if (ps[u]) ea[pj[u]] = pg[u];
    }
    qlo = wave_min(qlo);
    qhi = wave_max(qhi);
    if (lane == 0) {
      red[0][w] = qlo;
      red[1][w] = qhi;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kW; ++q) {
      qlo = min(qlo, red[0][q]);
      qhi = max(qhi, red[1][q]);
    }
    const uint32_t qmin = __builtin_amdgcn_readfirstlane(qlo);
    const uint32_t qmax = __builtin_amdgcn_readfirstlane(qhi);
    // keys relative to the chunk's smallest q20: only the bits of qmax - qmin are ranked
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) ka[r * kT + t] -= qmin << 12;
    const uint32_t span = qmax - qmin;
    lds_sort_span<kT, kRow>(ka, kN, span ? 32u - (uint32_t)__clz(span) : 0u, wc, wsum);
    uint32_t* dst = b.first + (size_t)l * kN;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t i = r * kT + t;
      dst[i] = ea[ka[i] & 0xFFFu];
    }
    __syncthreads();  // the next item reuses the shared arrays
  }
}

// One wave sorts a small tile of at most kWaveSortMax entries in place (entries and keys at off)
// by (q20, list index), as sort_segment does, with no workgroup barrier.  The keys ping-pong
// between ka[0, 512) and ka[512, 1024), and each key's rank in its digit waits in ka[1024, 1536)
// (few registers: the kernel's occupancy is set by its other items); each LSD pass ranks the keys
// with ballots against the wave's own 128 digit counters, scans those across the wave and
// scatters; ea holds the entries by list index.  A wave's LDS
// operations execute in issue order, so a wave-scope fence between the phases (which keeps the
// compiler from moving memory operations across it) is all the ordering they need.
constexpr uint32_t kWaveSortMax = 512;
__device__ __forceinline__ void wave_sort_small(const Bufs& b, uint32_t off, uint32_t cnt, uint32_t* ka,
                                                uint32_t* ea, uint32_t* wc) {
  const uint32_t lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t rounds = (cnt + 63u) / 64u;  // uniform, <= 8
  uint32_t qmin = 0xFFFFFFFFu, qmax = 0u;
  for (uint32_t i = lane; i < cnt; i += 64u) {
    const uint32_t q = b.qkey[off + i];
    ea[i] = b.entries[off + i];
    ka[i] = q;
    qmin = min(qmin, q);
    qmax = max(qmax, q);
  }
  qmin = wave_min(qmin);
  qmax = wave_max(qmax);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (uint32_t i = lane; i < cnt; i += 64u) ka[i] = ((ka[i] - qmin) << 12) | i;
  const uint32_t span = qmax - qmin;  // uniform
  const uint32_t qbits = span ? 32u - (uint32_t)__clz(span) : 0u;
  const uint32_t np = (qbits + 6u) / 7u;
  const uint32_t w0 = np ? max((qbits + np - 1u) / np, 4u) : 0u;
  uint32_t* src = ka;
  uint32_t* dst = ka + kWaveSortMax;
  uint32_t* const rk = ka + 2u * kWaveSortMax;
  uint32_t low = 0;
  for (uint32_t ps = 0; ps < np; ++ps) {  // uniform
    const uint32_t bits = ps + 1u < np ? w0 : max(qbits - low, 4u);
    const uint32_t shift = 12u + low, mask = (1u << bits) - 1u;
    wc[lane] = 0u;
    wc[lane + 64u] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll 1
    for (uint32_t r = 0; r < rounds; ++r) {
      const bool valid = r * 64u + lane < cnt;
      const uint32_t d = valid ? (src[r * 64u + lane] >> shift) & mask : 0u;  // < 128
      const unsigned long long m = match_digit<7>(d, __ballot(valid));
      const uint32_t old = group_count(&wc[d], m, valid, lane);
      if (valid) rk[r * 64u + lane] = old + (uint32_t)__popcll(m & below);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // exclusive scan of the 128 digit counts: digit lane, then digit lane + 64
    const uint32_t c0 = wc[lane], c1 = wc[lane + 64u];
    uint32_t i0 = c0, i1 = c1;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
      const uint32_t x0 = __shfl_up(i0, dd, 64), x1 = __shfl_up(i1, dd, 64);
      if (lane >= (uint32_t)dd) {
        i0 += x0;
        i1 += x1;
      }
    }
    const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)i0, 63);
    wc[lane] = i0 - c0;
    wc[lane + 64u] = t0 + i1 - c1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t i = lane; i < cnt; i += 64u) {
      const uint32_t key = src[i];
      dst[wc[(key >> shift) & mask] + rk[i]] = key;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t* const tmp = src;
    src = dst;
    dst = tmp;
    low += bits;
  }
  for (uint32_t i = lane; i < cnt; i += 64u) {
    const uint32_t key = src[i];
    b.entries[off + i] = ea[key & 0xFFFu];
    if (b.keys20) b.keys20[off + i] = key + (qmin << 12);
  }
}

// Round 1 of a deferred frame with first chunks: every large tile's first chunk (largest tiles
// first, from large_order[]) and then every small tile's sort, one workgroup per item, in one
// launch: the small tiles fill the CUs the last large tiles leave idle.  A small tile of at most
// kWaveSortMax entries is sorted by the workgroup's first wave alone (wave_sort_small).
template <uint32_t kT>
__global__ __launch_bounds__(kT, 8) void k_first_chunks(Bufs b) {
  __shared__ FirstLds<kT> S_;
  const uint32_t nl = min(b.stats->large_tiles, b.first_tiles);
  const uint32_t n = nl + b.stats->small_tiles;
  for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
    if (k < nl) {
      first_chunk_tile<kT>(b, __builtin_amdgcn_readfirstlane(b.large_order[k]), S_);
    } else {
      const uint32_t m = __builtin_amdgcn_readfirstlane(b.small_list[k - nl]);
      const uint32_t off = __builtin_amdgcn_readfirstlane(b.tile_offsets[m]);
      const uint32_t cnt = __builtin_amdgcn_readfirstlane(b.tile_counts[m]);
      if (cnt <= kWaveSortMax) {  // uniform
        if (threadIdx.x < 64) wave_sort_small(b, off, cnt, S_.ka, S_.ea, S_.wc[0]);
        __syncthreads();  // the shared arrays are reused by the next item
        continue;
      }
      sort_segment<kT, kFirstRow>(b, b.entries + off, b.qkey + off, cnt, off, S_.ka, S_.wc, S_.wsum, S_.red);
    }
  }
}

// Round 2 of a deferred frame, after the sort of the resumed tiles' later chunks: each of their
// entries' reach_quadrants mask in its tile, into qkey beside it (those keys are dead once the
// chunks are sorted), for the entries before HDSR_RESUME_MASKED in the tile's list.  The resumed
// rasterizer then gathers only the records of the entries that may reach its quadrant; past that
// bound (a tile that stays unsaturated that long) it tests the records itself, so a frame with
// very long resumed tiles does not test entries its pixels never reach.  One thread per entry:
// item (segment k, 256-entry part j).
__global__ __launch_bounds__(256) void k_resume_masks(Bufs b) {
  const uint32_t n = b.stats->pend_segs * (HDSR_MAX_TILE_ENTRIES / 256u);
  for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
    const Seg S = b.segs[b.pend_segs[it / (HDSR_MAX_TILE_ENTRIES / 256u)]];
    const uint32_t i = (it % (HDSR_MAX_TILE_ENTRIES / 256u)) * 256u + threadIdx.x;
    if (i >= S.count) continue;
    const LargeTile& L = b.large_tiles[S.large];
    if (S.dst - L.start >= HDSR_RESUME_MASKED) continue;  // uniform: the segment's first entry
    const uint32_t t = L.tile;
    const FrameDesc& F = b.frames[frame_of(b.tile_frame_start, b.num_frames, t)];
    const uint32_t lt = t - F.tile_base;
    const float x0 = (float)((lt % F.tiles_x) * 16u) + 0.5f, y0 = (float)((lt / F.tiles_x) * 16u) + 0.5f;
    const uint32_t g = b.entries[S.dst + i];
    const uint4 h0 = b.hot[(size_t)g * 2];
    const float k2 = hdsr_u2f(b.hot[(size_t)g * 2 + 1].x);
    b.qkey[S.dst + i] = reach_quadrants(hdsr_u2f(h0.x), hdsr_u2f(h0.y), hdsr_u2f(h0.z), hdsr_u2f(h0.w), k2, x0, y0);
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
  if (b.round == 2) {  // the few tiles the first round left unsaturated, the longest ones: wide
    const uint32_t grid = b.total_tiles < 256u ? b.total_tiles : 256u;
    hipLaunchKernelGGL(k_bucket<1024>, dim3(grid), dim3(1024), 0, stream, b);
    return;
  }
  if (b.bucket_wide) {
    const uint32_t grid = b.total_tiles < 1024u ? b.total_tiles : 1024u;
    hipLaunchKernelGGL(k_bucket<1024>, dim3(grid), dim3(1024), 0, stream, b);
  } else {
    const uint32_t grid = b.total_tiles < 4096u ? b.total_tiles : 4096u;
    hipLaunchKernelGGL(k_bucket<512>, dim3(grid), dim3(512), 0, stream, b);
  }
}

void launch_first_chunks(const Bufs& b, hipStream_t stream) {
  if (!b.total_tiles) return;
  hipLaunchKernelGGL(k_order_large, dim3(1), dim3(1024), 0, stream, b);
  // one workgroup per item (large and small tiles together never exceed the tiles)
  const uint32_t grid = b.total_tiles;
  if (b.bucket_wide)
    hipLaunchKernelGGL(k_first_chunks<1024>, dim3(grid), dim3(1024), 0, stream, b);
  else
    hipLaunchKernelGGL(k_first_chunks<512>, dim3(grid), dim3(512), 0, stream, b);
}

void launch_segment_sort(const Bufs& b, hipStream_t stream) {
  if (!b.seg_cap) return;
  // round 2 sorts the listed chunks only: a grid that holds them at once without dispatching
  // thousands of workgroups that find nothing to do
  const uint32_t cap = b.round == 2 ? 1024u : 2048u * 1024u / kSortThreads;
  const uint32_t grid = b.seg_cap < cap ? b.seg_cap : cap;
  if (b.round == 2 && b.first_chunks)  // the frame selected first chunks: few tiles resume
    hipLaunchKernelGGL(k_segsort<1024>, dim3(grid), dim3(1024), 0, stream, b);
  else
    hipLaunchKernelGGL(k_segsort<kSortThreads>, dim3(grid), dim3(kSortThreads), 0, stream, b);
  if (b.round == 2)  // the resumed entries' quadrant masks
    hipLaunchKernelGGL(k_resume_masks, dim3(4096), dim3(256), 0, stream, b);
  if (b.ref_caps) hipLaunchKernelGGL(k_segcopy, dim3(grid), dim3(256), 0, stream, b);
}

}  // namespace hdsr
