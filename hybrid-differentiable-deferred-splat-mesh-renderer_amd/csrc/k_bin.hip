// k_bin.hip - count -> hierarchical prefix scan -> scatter -> radix grouping -> tile ranges -> classify.
//
// Reference: count_tile_refs / scatter_tile_refs (src/shaders/gpu_tile_splat_composite.wgsl:64-119),
// large scatter (gpu_large_splat_scatter.wgsl:52-89), scan / scan_block_sums / add_block_sums
// (gpu_prefix_sum.wgsl:14-85), the 4-bit LSD radix sort (gpu_radix_sort.wgsl:27-136,
// src/gpuRadixSort.ts:313-354) and classify_tiles (gpu_tile_classify.wgsl:44-86).
//
// The reference resolves slots with atomics, so its per-tile lists come out in arbitrary order.
// Here the order is canonical by construction: every visible splat's pairs (Morton key, rank)
// are written at an offset from an exclusive scan over the per-splat tile counts (small splats
// in ascending rank, then large splats in ascending rank), and a stable LSD radix sort on the
// Morton key groups them by tile while keeping that order.  The key width decides the passes
// (at most 8 bits each: a 1080p frame's 14-bit Morton range takes two 7-bit passes, as the
// reference's 4 x 4-bit passes over 16 bits give the same stable order); each pass is a
// per-workgroup digit histogram, a hierarchical scan of the digit-major histogram, and a stable
// scatter that ranks equal digits inside each wave with ballots.  All of it is deterministic.
#include "hdsr_internal.h"

namespace hdsr {
namespace {

__device__ __forceinline__ uint4 operator+(uint4 a, uint4 b) {
  return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

template <typename T>
struct Ops;
template <>
struct Ops<uint32_t> {
  __device__ static uint32_t zero() { return 0u; }
  __device__ static uint32_t up(uint32_t v, int d) { return __shfl_up(v, d, 64); }
  __device__ static uint32_t add(uint32_t a, uint32_t b) { return a + b; }
};
template <>
struct Ops<uint4> {
  __device__ static uint4 zero() { return make_uint4(0, 0, 0, 0); }
  __device__ static uint4 up(uint4 v, int d) {
    return make_uint4(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64), __shfl_up(v.z, d, 64), __shfl_up(v.w, d, 64));
  }
  __device__ static uint4 add(uint4 a, uint4 b) { return a + b; }
};

// Exclusive scan over a 256-thread workgroup of one value per thread; *total gets the sum.
template <typename T>
__device__ __forceinline__ T block_excl_256(T v, T* total) {
  __shared__ T wsum[4];
  const uint32_t lane = __lane_id();
  T inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T t = Ops<T>::up(inc, d);
    if (lane >= (uint32_t)d) inc = Ops<T>::add(inc, t);
  }
  T exc = Ops<T>::up(inc, 1);
  if (lane == 0) exc = Ops<T>::zero();
  const uint32_t w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  T base = Ops<T>::zero(), tot = Ops<T>::zero();
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    const T s = wsum[i];
    if (i < w) base = Ops<T>::add(base, s);
    tot = Ops<T>::add(tot, s);
  }
  __syncthreads();
  *total = tot;
  return Ops<T>::add(base, exc);
}

// ---- generic hierarchical exclusive scan: 4096 elements per workgroup (256 threads x 16) ----

template <typename T>
__global__ __launch_bounds__(256) void k_scan_reduce(const T* in, T* partial, uint32_t n) {
  const uint32_t base = blockIdx.x * kScanChunk;
  T s = Ops<T>::zero();
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    const uint32_t i = base + j * 256 + threadIdx.x;
    if (i < n) s = Ops<T>::add(s, in[i]);
  }
  T tot;
  (void)block_excl_256<T>(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Exclusive scan of one chunk (thread t owns elements 16t..16t+15), plus a carried offset.
template <typename T>
__global__ __launch_bounds__(256) void k_scan_down(const T* in, T* out, uint32_t n, const T* offsets) {
  const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * 16;
  T v[16];
  T s = Ops<T>::zero();
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    v[j] = (base + j < n) ? in[base + j] : Ops<T>::zero();
    s = Ops<T>::add(s, v[j]);
  }
  T tot;
  T run = block_excl_256<T>(s, &tot);
  if (offsets) run = Ops<T>::add(run, offsets[blockIdx.x]);
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    if (base + j < n) out[base + j] = run;
    run = Ops<T>::add(run, v[j]);
  }
}

template <typename T>
size_t scan_tmp_elems(uint32_t n) {
  size_t total = 0;
  while (n > kScanChunk) {
    const uint32_t parts = (n + kScanChunk - 1) / kScanChunk;
    total += 2 * (size_t)parts;
    n = parts;
  }
  return total + 1;
}

template <typename T>
void scan_exclusive(const T* in, T* out, uint32_t n, T* tmp, hipStream_t stream) {
  if (n == 0) return;
  const uint32_t parts = (n + kScanChunk - 1) / kScanChunk;
  if (parts == 1) {
    hipLaunchKernelGGL(k_scan_down<T>, dim3(1), dim3(256), 0, stream, in, out, n, (const T*)nullptr);
    return;
  }
  T* partial = tmp;
  T* partial_scan = tmp + parts;
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3(parts), dim3(256), 0, stream, in, partial, n);
  scan_exclusive<T>(partial, partial_scan, parts, tmp + 2 * (size_t)parts, stream);
  hipLaunchKernelGGL(k_scan_down<T>, dim3(parts), dim3(256), 0, stream, in, out, n, (const T*)partial_scan);
}

// ---- per-splat counts -> pair offsets -------------------------------------------------------

__device__ __forceinline__ void box_of(uint32_t bd, uint32_t* x0, uint32_t* y0, uint32_t* x1, uint32_t* y1) {
  *x0 = bd & 0xFFu;
  *y0 = (bd >> 8) & 0xFFu;
  *x1 = (bd >> 16) & 0xFFu;
  *y1 = bd >> 24;
}

// The four per-splat counts (small pairs, large pairs, is large, visible) from its tile box.
__device__ __forceinline__ uint4 counts_of(uint32_t bd) {
  if (bd == HDSR_SENTINEL) return make_uint4(0, 0, 0, 0);
  uint32_t x0, y0, x1, y1;
  box_of(bd, &x0, &y0, &x1, &y1);
  const uint32_t span = (x1 - x0 + 1u) * (y1 - y0 + 1u);
  const uint32_t nt = (x1 >= x0 && y1 >= y0) ? (x1 - x0 + 1u) * (y1 - y0 + 1u) : 0u;
  if (span > HDSR_LARGE_SPLAT_TILES) return make_uint4(0, nt, 1, 1);
  return make_uint4(nt, 0, 0, 1);
}

__global__ void k_clear_stats(DevStats* s) {
  uint32_t* p = reinterpret_cast<uint32_t*>(s);
  if (threadIdx.x < sizeof(DevStats) / 4) p[threadIdx.x] = 0;
}

__global__ void k_totals(Bufs b) {
  const uint32_t last = b.splat_blocks - 1;
  const uint4 t = b.block_pre[last] + b.block_cnt[last];
  DevStats* s = b.stats;
  s->small_pairs = t.x;
  s->large_pairs = t.y;
  s->large_splats = t.z;
  s->visible = t.w;
  const uint64_t refs = (uint64_t)t.x + t.y;
  s->refs = refs > b.cap ? b.cap : (uint32_t)refs;
  s->overflow = refs > b.cap ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_emit(Bufs b) {
  const uint32_t blk = blockIdx.x;
  const uint32_t f = frame_of(b.block_frame_start, b.num_frames, blk);
  const uint32_t mbase = b.frames[f].morton_base;
  const uint32_t g = blk * kSplatBlock + threadIdx.x;
  const uint32_t bd = b.bounds[g];
  const uint4 c = counts_of(bd);
  uint4 tot;
  const uint4 pre = block_excl_256<uint4>(c, &tot) + b.block_pre[blk];
  const uint32_t small_total = b.stats->small_pairs;
  if (c.z) b.large_list[pre.z] = g;
  // A wave's small-splat pairs occupy one contiguous range of the output and its large-splat
  // pairs another (both classes are in ascending rank), so the wave writes each range with
  // consecutive lanes: pair p of the range belongs to the lane whose running start is the last
  // one <= p, and is that splat's (p - start)-th tile in row-major order of its box.
  const uint32_t lane = __lane_id();
  const uint32_t cap = b.cap;
  const uint32_t g0 = g - lane;
#pragma unroll
  for (uint32_t cls = 0; cls < 2; ++cls) {
    const uint32_t n = cls ? c.y : c.x;
    uint32_t inc = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t x = __shfl_up(inc, d, 64);
      if (lane >= (uint32_t)d) inc += x;
    }
    const uint32_t loc = inc - n;
    const uint32_t total = __shfl(inc, 63, 64);
    const uint32_t off = cls ? small_total + pre.y : pre.x;
    for (uint32_t p0 = 0; p0 < total; p0 += 64) {
      const uint32_t p = p0 + lane;
      uint32_t lo = 0, hi = 64;
#pragma unroll
      for (int it = 0; it < 6; ++it) {
        const uint32_t mid = (lo + hi) >> 1;
        if (__shfl(loc, mid, 64) <= p) lo = mid;
        else hi = mid;
      }
      const uint32_t o_loc = __shfl(loc, lo, 64);
      const uint32_t o_off = __shfl(off, lo, 64);
      const uint32_t o_bd = __shfl(bd, lo, 64);
      if (p < total) {
        const uint32_t j = p - o_loc;
        uint32_t x0, y0, x1, y1;
        box_of(o_bd, &x0, &y0, &x1, &y1);
        const uint32_t sx = x1 - x0 + 1u;
        const uint32_t ty = y0 + j / sx, tx = x0 + j % sx;
        const uint32_t dst = o_off + j;
        if (dst < cap) {
          b.keys_a[dst] = mbase + hdsr_morton2(tx, ty);
          b.vals_a[dst] = g0 + lo;
        }
      }
    }
  }
}

// ---- stable LSD radix passes on the Morton key ---------------------------------------------
//
// A pass over blocks of kRadixBlockItems pairs: k_radix_hist counts the digit of each key per
// block (digit-major table, then scanned), k_radix_scatter ranks each pair inside its block and
// writes it to its global slot.  In the scatter each wave owns 1024 consecutive pairs of the
// block and ranks equal digits with ballots plus a wave-private running count (no workgroup
// barrier per round); the block's pairs are then reordered in LDS so that the global writes of
// each digit run are contiguous.

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, uint32_t* hist, const DevStats* st,
                                                    uint32_t nblocks, uint32_t shift, uint32_t mask) {
  __shared__ uint32_t cnt[kRadixBuckets];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = st->refs;
  const uint32_t base = blockIdx.x * kRadixBlockItems;
  if (base < n) {
#pragma unroll 4
    for (uint32_t j = 0; j < 16; ++j) {
      const uint32_t i = base + j * 256 + threadIdx.x;
      if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & mask], 1u);
    }
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// With 'dlog' set (the last pass), also writes each moved pair's (depth bits, log-depth) to
// ekey_out at the same position, so the per-tile sorts read them as a stream.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                       uint32_t* vout, const uint32_t* hist, const DevStats* st,
                                                       uint32_t nblocks, uint32_t shift, uint32_t bits,
                                                       const uint2* dlog, uint2* ekey_out) {
  __shared__ uint32_t sk[kRadixBlockItems];
  __shared__ uint32_t sv[kRadixBlockItems];
  __shared__ uint32_t wc[4][kRadixBuckets];
  __shared__ uint32_t dbase[kRadixBuckets];
  __shared__ uint32_t gbase[kRadixBuckets];
  __shared__ uint32_t wsum[4];
  const uint32_t n = st->refs;
  const uint32_t base = blockIdx.x * kRadixBlockItems;
  if (base >= n) return;  // uniform per workgroup
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t mask = (1u << bits) - 1u;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) wc[q][t] = 0;
  gbase[t] = hist[t * nblocks + blockIdx.x];
  __syncthreads();
  uint32_t k[16], v[16], pos[16];
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {  // all loads first: 32 in flight per lane
    const uint32_t i = base + w * 1024 + r * 64 + lane;
    const bool valid = i < n;
    k[r] = valid ? kin[i] : 0u;
    v[r] = valid ? vin[i] : 0u;
  }
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {
    const uint32_t i = base + w * 1024 + r * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = (k[r] >> shift) & mask;
    unsigned long long m = __ballot(valid);
    for (uint32_t bit = 0; bit < bits; ++bit) {
      const bool s = (d >> bit) & 1u;
      const unsigned long long bm = __ballot(s);
      m &= s ? bm : ~bm;
    }
    const uint32_t rank = __popcll(m & below);
    uint32_t old = 0;
    if (valid) old = wc[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && (m >> lane) == 1ull) wc[w][d] = old + rank + 1;
    __builtin_amdgcn_wave_barrier();
    pos[r] = valid ? old + rank : 0xFFFFFFFFu;
  }
  __syncthreads();
  // block-local digit starts (exclusive scan over digits) and each wave's start inside a digit
  const uint32_t c0 = wc[0][t], c1 = wc[1][t], c2 = wc[2][t], c3 = wc[3][t];
  const uint32_t tot = c0 + c1 + c2 + c3;
  uint32_t inc = tot;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t x = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += x;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t db = inc - tot;
  for (uint32_t q = 0; q < w; ++q) db += wsum[q];
  dbase[t] = db;
  wc[0][t] = db;
  wc[1][t] = db + c0;
  wc[2][t] = db + c0 + c1;
  wc[3][t] = db + c0 + c1 + c2;
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {
    if (pos[r] != 0xFFFFFFFFu) {
      const uint32_t lp = wc[w][(k[r] >> shift) & mask] + pos[r];
      sk[lp] = k[r];
      sv[lp] = v[r];
    }
  }
  __syncthreads();
  const uint32_t cnt = min(kRadixBlockItems, n - base);
  for (uint32_t q = t; q < cnt; q += 256) {
    const uint32_t kk = sk[q];
    const uint32_t d = (kk >> shift) & mask;
    const uint32_t dst = gbase[d] + (q - dbase[d]);
    const uint32_t vv = sv[q];
    kout[dst] = kk;
    vout[dst] = vv;
    if (dlog) ekey_out[dst] = dlog[vv];
  }
}

// ---- tile ranges: offsets / counts per global Morton slot ---------------------------------

__global__ void k_tile_ranges(const uint32_t* tkeys, const DevStats* st, uint32_t total_morton,
                              uint32_t* offsets, uint32_t* counts) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= total_morton) return;
  const uint32_t n = st->refs;
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tkeys[mid] < m) lo = mid + 1; else hi = mid;
  }
  const uint32_t start = lo;
  hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tkeys[mid] <= m) lo = mid + 1; else hi = mid;
  }
  offsets[m] = start;
  counts[m] = lo - start;
}

// ---- classify: small / large tile lists and the sort segments, in row-major tile order -----

__device__ __forceinline__ void tile_info(const Bufs& b, uint32_t t, uint32_t* morton, uint32_t* start,
                                          uint32_t* count) {
  const uint32_t f = frame_of(b.tile_frame_start, b.num_frames, t);
  const FrameDesc& F = b.frames[f];
  const uint32_t lt = t - F.tile_base;
  const uint32_t m = F.morton_base + hdsr_morton2(lt % F.tiles_x, lt / F.tiles_x);
  *morton = m;
  *start = b.tile_offsets[m];
  *count = b.tile_counts[m];
}

// (is small, is large, segments, scratch entries) of one tile
__device__ __forceinline__ uint4 class_of(uint32_t c) {
  if (c == 0) return make_uint4(0, 0, 0, 0);
  if (c <= HDSR_MAX_TILE_ENTRIES) return make_uint4(1, 0, 1, 0);
  return make_uint4(0, 1, (c + HDSR_MAX_TILE_ENTRIES - 1) / HDSR_MAX_TILE_ENTRIES, c);
}

__global__ void k_classify_count(Bufs b, uint4* cls) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.total_tiles) return;
  uint32_t m, s, c;
  tile_info(b, t, &m, &s, &c);
  cls[t] = class_of(c);
}

__global__ void k_classify_write(Bufs b, const uint4* pre) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.total_tiles) return;
  uint32_t m, s, c;
  tile_info(b, t, &m, &s, &c);
  const uint4 k = class_of(c);
  const uint4 p = pre[t];
  if (t == b.total_tiles - 1) {
    b.stats->small_tiles = p.x + k.x;
    b.stats->large_tiles = p.y + k.y;
    b.stats->segments = p.z + k.z;
    b.stats->chunks = p.z + k.z - (p.x + k.x);
    b.stats->scratch_used = p.w + k.w;
  }
  if (k.x) {
    b.small_list[p.x] = m;
    b.segs[p.z] = Seg{s, s, c, 0xFFFFFFFFu};
  } else if (k.y) {
    b.large_list_t[p.y] = m;
    b.large_bases[p.y] = b.stats->refs + p.w;
    b.large_tiles[p.y] = LargeTile{m, s, c, p.w, 0u, 0u, 0u, 0u};
    for (uint32_t q = 0; q < k.z; ++q) {
      const uint32_t off = q * HDSR_MAX_TILE_ENTRIES;
      const uint32_t cnt = min(HDSR_MAX_TILE_ENTRIES, c - off);
      b.segs[p.z + q] = Seg{p.w + off, s + off, cnt, p.y};
    }
  }
}

}  // namespace

size_t scan_u32_tmp_words(uint32_t n) { return scan_tmp_elems<uint32_t>(n); }
size_t scan_u4_tmp_words(uint32_t n) { return scan_tmp_elems<uint4>(n); }
void scan_u32_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t stream) {
  scan_exclusive<uint32_t>(in, out, n, tmp, stream);
}
void scan_u4_exclusive(const uint4* in, uint4* out, uint32_t n, uint4* tmp, hipStream_t stream) {
  scan_exclusive<uint4>(in, out, n, tmp, stream);
}

void launch_clear_stats(const Bufs& b, hipStream_t stream) {
  hipLaunchKernelGGL(k_clear_stats, dim3(1), dim3(64), 0, stream, b.stats);
}

void launch_scan_blocks(const Bufs& b, hipStream_t stream) {
  if (!b.splat_blocks) return;
  scan_u4_exclusive(b.block_cnt, b.block_pre, b.splat_blocks, b.scan_tmp, stream);
  hipLaunchKernelGGL(k_totals, dim3(1), dim3(1), 0, stream, b);
}

void launch_emit_pairs(const Bufs& b, hipStream_t stream) {
  if (!b.splat_blocks) return;
  hipLaunchKernelGGL(k_emit, dim3(b.splat_blocks), dim3(kSplatBlock), 0, stream, b);
}

void launch_group_pairs(const Bufs& b, hipStream_t stream) {
  const uint32_t nblocks = (b.cap + kRadixBlockItems - 1) / kRadixBlockItems;
  if (!nblocks) return;
  const uint32_t passes = b.key_bits == 0 ? 1u : (b.key_bits + kRadixBits - 1) / kRadixBits;
  const uint32_t bits = b.key_bits == 0 ? 1u : (b.key_bits + passes - 1) / passes;
  const uint32_t mask = (1u << bits) - 1u;
  const uint32_t* ks = b.keys_a;
  const uint32_t* vs = b.vals_a;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = p * bits;
    const bool last = p + 1 == passes;
    // the last pass writes the grouped entries and their keys; earlier ones ping-pong
    uint32_t* kd = last ? b.tkeys : (ks == b.keys_a ? b.keys_b : b.keys_a);
    uint32_t* vd = last ? b.entries : (vs == b.vals_a ? b.vals_b : b.vals_a);
    hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(256), 0, stream, ks, b.radix_hist, b.stats, nblocks, shift,
                       mask);
    scan_u32_exclusive(b.radix_hist, b.radix_hist, nblocks * kRadixBuckets, b.radix_tmp, stream);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(256), 0, stream, ks, vs, kd, vd, b.radix_hist,
                       b.stats, nblocks, shift, bits, last ? (const uint2*)b.dlog : (const uint2*)nullptr,
                       last ? b.ekey : (uint2*)nullptr);
    ks = kd;
    vs = vd;
  }
}

void launch_tile_ranges(const Bufs& b, hipStream_t stream) {
  if (!b.total_morton) return;
  hipLaunchKernelGGL(k_tile_ranges, dim3((b.total_morton + 255) / 256), dim3(256), 0, stream, b.tkeys, b.stats,
                     b.total_morton, b.tile_offsets, b.tile_counts);
}

void launch_classify(const Bufs& b, hipStream_t stream) {
  if (!b.total_tiles) return;
  // the class vector and its prefix live in the uint4 scan area after the block prefix scratch
  uint4* cls = b.scan_tmp + scan_u4_tmp_words(b.splat_blocks);
  uint4* pre = cls + b.total_tiles;
  uint4* tmp = pre + b.total_tiles;
  const uint32_t grid = (b.total_tiles + 255) / 256;
  hipLaunchKernelGGL(k_classify_count, dim3(grid), dim3(256), 0, stream, b, cls);
  scan_u4_exclusive(cls, pre, b.total_tiles, tmp, stream);
  hipLaunchKernelGGL(k_classify_write, dim3(grid), dim3(256), 0, stream, b, (const uint4*)pre);
}

}  // namespace hdsr
