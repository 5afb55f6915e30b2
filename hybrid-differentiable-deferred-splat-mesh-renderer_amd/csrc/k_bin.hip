// k_bin.hip - count -> hierarchical prefix scan -> pair emission -> grouping into tile lists ->
// classify.
//
// Reference: count_tile_refs / scatter_tile_refs (src/shaders/gpu_tile_splat_composite.wgsl:64-119),
// large scatter (gpu_large_splat_scatter.wgsl:52-89), scan / scan_block_sums / add_block_sums
// (gpu_prefix_sum.wgsl:14-85), the 4-bit LSD radix sort (gpu_radix_sort.wgsl:27-136,
// src/gpuRadixSort.ts:313-354) and classify_tiles (gpu_tile_classify.wgsl:44-86).
//
// The reference resolves slots with atomics, so its per-tile lists come out in arbitrary order.
// Here the order is canonical by construction: every tile's list holds the splats whose boxes
// cover it, small splats in ascending rank, then large splats in ascending rank (the order a
// stable sort of the reference's pairs on the Morton key gives).  The lists are built in two
// levels.  Each visible splat emits one (cell, rank) pair per 8 x 8-tile cell its box touches
// (about 3 per splat instead of about 40 tile pairs), at offsets from an exclusive scan of the
// per-workgroup counts; a stable LSD radix sort on the global cell id groups them (one 8-bit pass
// for a 1080p frame's 135 cells); then every cell's list is expanded into its 64 tiles' lists
// with ballots (counting pass, per-cell prefix over chunks, scan of the tile counts over Morton
// slots, writing pass).  All of it is deterministic.
#include <algorithm>

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
  __device__ static uint32_t sub(uint32_t a, uint32_t b) { return a - b; }
};
template <>
struct Ops<uint4> {
  __device__ static uint4 zero() { return make_uint4(0, 0, 0, 0); }
  __device__ static uint4 up(uint4 v, int d) {
    return make_uint4(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64), __shfl_up(v.z, d, 64), __shfl_up(v.w, d, 64));
  }
  __device__ static uint4 add(uint4 a, uint4 b) { return a + b; }
  __device__ static uint4 sub(uint4 a, uint4 b) { return make_uint4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
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

// One 1024-thread workgroup scans up to kOneGroupScan elements: the tile-count scan of a frame
// then takes one launch instead of the hierarchical path's three.
constexpr uint32_t kOneGroupScan = 16384;

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

// One thread: clears the statistics (no earlier stage of the frame writes them) and stores the
// totals of the block scans.
__global__ void k_totals(Bufs b) {
  uint32_t* st = reinterpret_cast<uint32_t*>(b.stats);
  for (uint32_t i = 0; i < sizeof(DevStats) / 4; ++i) st[i] = 0;
  const uint32_t last = b.splat_blocks - 1;
  const uint4 t = b.block_pre[last] + b.block_cnt[last];
  const uint4 c = b.block_ccpre[last] + b.block_cc[last];
  DevStats* s = b.stats;
  s->small_pairs = t.x;
  s->large_pairs = t.y;
  s->large_splats = t.z;
  s->visible = t.w;
  const uint64_t refs = (uint64_t)t.x + t.y;
  s->refs = refs > b.cap ? b.cap : (uint32_t)refs;
  s->overflow = (!b.ref_caps && refs > b.cap) ? 1u : 0u;  // under the caps, dropping is the result
  const uint64_t cells = (uint64_t)c.x + c.y;  // <= kMaxCellsPerSplat per splat by construction
  s->cell_pairs = cells > b.cell_cap ? b.cell_cap : (uint32_t)cells;
}

// Both block-count scans and the totals in one 1024-thread workgroup (splat_blocks <= 4096):
// thread t holds blocks 4t .. 4t+3 of both arrays in registers, loaded at once; the thread that
// holds the last block writes the totals as k_totals does.
constexpr uint32_t kOneGroupBlocks = 4096;
__global__ __launch_bounds__(1024) void k_scan_blocks_one(Bufs b) {
  __shared__ uint4 wsum[2][16];
  const uint32_t n = b.splat_blocks;
  const uint32_t t = threadIdx.x, lane = __lane_id(), w = t >> 6, i0 = t * 4u;
  uint4 a[4], c[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    a[k] = i0 + k < n ? b.block_cnt[i0 + k] : make_uint4(0, 0, 0, 0);
    c[k] = i0 + k < n ? b.block_cc[i0 + k] : make_uint4(0, 0, 0, 0);
  }
  const uint4 pa = a[0] + a[1] + a[2] + a[3], pc = c[0] + c[1] + c[2] + c[3];
  uint4 ia = pa, ic = pc;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {  // inclusive scans across the wave
    const uint4 xa = Ops<uint4>::up(ia, d), xc = Ops<uint4>::up(ic, d);
    if (lane >= (uint32_t)d) {
      ia = ia + xa;
      ic = ic + xc;
    }
  }
  if (lane == 63) {
    wsum[0][w] = ia;
    wsum[1][w] = ic;
  }
  __syncthreads();
  uint4 ra = Ops<uint4>::sub(ia, pa), rc = Ops<uint4>::sub(ic, pc);
#pragma unroll
  for (uint32_t q = 0; q < 16; ++q)
    if (q < w) {
      ra = ra + wsum[0][q];
      rc = rc + wsum[1][q];
    }
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    if (i0 + k < n) {
      b.block_pre[i0 + k] = ra;
      b.block_ccpre[i0 + k] = rc;
    }
    if (i0 + k == n - 1u) {  // the totals (one thread)
      const uint4 tt = ra + a[k], cc = rc + c[k];
      uint32_t* st = reinterpret_cast<uint32_t*>(b.stats);
      for (uint32_t j = 0; j < sizeof(DevStats) / 4; ++j) st[j] = 0;
      DevStats* sp = b.stats;
      sp->small_pairs = tt.x;
      sp->large_pairs = tt.y;
      sp->large_splats = tt.z;
      sp->visible = tt.w;
      const uint64_t refs = (uint64_t)tt.x + tt.y;
      sp->refs = refs > b.cap ? b.cap : (uint32_t)refs;
      sp->overflow = (!b.ref_caps && refs > b.cap) ? 1u : 0u;
      const uint64_t cells = (uint64_t)cc.x + cc.y;
      sp->cell_pairs = cells > b.cell_cap ? b.cell_cap : (uint32_t)cells;
    }
    ra = ra + a[k];
    rc = rc + c[k];
  }
}

// The cell box of a tile box: first cell, width and height in cells (0 when invisible / empty).
__device__ __forceinline__ uint32_t cell_box(uint32_t bd, uint32_t* cx0, uint32_t* cy0, uint32_t* cw) {
  if (bd == HDSR_SENTINEL) return 0u;
  uint32_t x0, y0, x1, y1;
  box_of(bd, &x0, &y0, &x1, &y1);
  if (x1 < x0 || y1 < y0) return 0u;
  *cx0 = x0 >> 3;
  *cy0 = y0 >> 3;
  *cw = (x1 >> 3) - *cx0 + 1u;
  return *cw * ((y1 >> 3) - *cy0 + 1u);
}

// First level of the tile lists: one (global cell id, g) pair per cell a visible splat's box
// touches, small splats' pairs first (ascending rank), then large splats' (ascending rank).  A
// wave's pairs of one class are one contiguous output range, written with consecutive lanes:
// pair p belongs to the lane whose running start is the last one <= p.  Also writes the list of
// large splats (classify_large_splats, gpu_classify_large_splats.wgsl:33-56).
__global__ __launch_bounds__(256) void k_emit(Bufs b) {
  const uint32_t blk = blockIdx.x;
  const uint32_t f = frame_of(b.block_frame_start, b.num_frames, blk);
  const FrameDesc& F = b.frames[f];
  const uint32_t g = blk * kSplatBlock + threadIdx.x;
  const uint32_t bd = b.bounds[g];
  const uint4 c = counts_of(bd);
  uint32_t cx0 = 0, cy0 = 0, cw = 1;
  const uint32_t nc = cell_box(bd, &cx0, &cy0, &cw);
  const uint4 v = make_uint4(c.z ? 0u : nc, c.z ? nc : 0u, c.z, 0u);
  uint4 tot;
  const uint4 bp = b.block_ccpre[blk];
  const uint4 pre = block_excl_256<uint4>(v, &tot) + make_uint4(bp.x, bp.y, b.block_pre[blk].z, 0u);
  if (c.z && pre.z < b.splat_blocks * kSplatBlock) b.large_list[pre.z] = g;
  const uint32_t small_total = b.block_ccpre[b.splat_blocks - 1].x + b.block_cc[b.splat_blocks - 1].x;
  const uint32_t lane = __lane_id();
  const uint32_t g0 = g - lane;
#pragma unroll
  for (uint32_t cls = 0; cls < 2; ++cls) {
    const uint32_t n = cls ? v.y : v.x;
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
      const uint32_t o_x = __shfl(cx0, lo, 64), o_y = __shfl(cy0, lo, 64), o_w = __shfl(cw, lo, 64);
      if (p < total) {
        const uint32_t j = p - o_loc;
        const uint32_t dst = o_off + j;
        if (dst < b.cell_cap) {
          b.ckeys_a[dst] = F.cell_base + (o_y + j / o_w) * F.cells_x + (o_x + j % o_w);
          b.cvals_a[dst] = g0 + lo;
        }
      }
    }
  }
}

// ---- stable LSD radix passes on the cell id -------------------------------------------------
//
// A pass over blocks of kRadixBlockItems pairs: k_radix_hist counts the digit of each key per
// block (digit-major table, then scanned), k_radix_scatter ranks each pair inside its block and
// writes it to its global slot.  In the scatter each wave owns 1024 consecutive pairs of the
// block and ranks equal digits with ballots plus a wave-private running count (no workgroup
// barrier per round); the block's pairs are then reordered in LDS so that the global writes of
// each digit run are contiguous.  *count is the number of pairs (device).

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, uint32_t* hist, const uint32_t* count,
                                                    uint32_t nblocks, uint32_t shift, uint32_t mask) {
  __shared__ uint32_t cnt[kRadixBuckets];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = *count;
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

__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                       uint32_t* vout, const uint32_t* hist, const uint32_t* count,
                                                       uint32_t nblocks, uint32_t shift, uint32_t bits) {
  __shared__ uint32_t sk[kRadixBlockItems];
  __shared__ uint32_t sv[kRadixBlockItems];
  __shared__ uint32_t wc[4][kRadixBuckets];
  __shared__ uint32_t dbase[kRadixBuckets];
  __shared__ uint32_t gbase[kRadixBuckets];
  __shared__ uint32_t wsum[4];
  const uint32_t n = *count;
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
    kout[dst] = kk;
    vout[dst] = sv[q];
  }
}

// ---- cells -> tile lists ----------------------------------------------------------------------
//
// Each cell's grouped list (canonical order) is cut into chunks of kCellChunk entries; one
// 256-thread workgroup per chunk, each wave walking 256 consecutive entries in rounds of 64.
// A lane holds one entry's 64-bit mask of the cell's tiles its box covers; for tile t, a ballot
// of bit t gives the lanes covering t, and their ranks among those lanes.  Counting pass: lane t
// accumulates tile t's count.  After a per-cell prefix over chunks and a scan of the tile counts
// over Morton slots, the writing pass repeats the walk and stores each (tile, entry) at the
// tile's offset + the chunk's prefix + the waves before + the rank: the canonical order of the
// cell list is kept inside every tile.

// Also resets the per-tile counts and depth ranges that the counting walk accumulates into; with
// no pairs to group (keys null: no splats in the batch) it does only that.  When the cell ids
// were sorted in one radix pass (at most 256 cells), each cell is a digit and its range comes
// from the pass's scanned digit-major histogram (hist, nblocks per digit); otherwise from two
// binary searches in the sorted keys.
__global__ void k_cell_ranges(Bufs b, const uint32_t* keys, const uint32_t* hist, uint32_t nblocks) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = c; i < b.total_morton; i += gridDim.x * blockDim.x) {
    b.tile_counts[i] = 0u;
    b.tile_dmin[i] = 0xFFFFFFFFu;
    b.tile_dmax[i] = 0u;
    if (b.tile_unl) b.tile_unl[i] = 0u;
  }
  if (hist) {
    // one radix pass: at most 256 cells, all in workgroup 0, which also scans their chunk counts
    // into the chunk bases (no separate scan launch)
    if (blockIdx.x != 0) return;  // uniform
    const uint32_t n = b.stats->cell_pairs;
    const bool in = c < b.total_cells;
    const uint32_t s = in ? hist[(size_t)c * nblocks] : 0u;
    const uint32_t e = !in ? 0u : c + 1u < kRadixBuckets ? hist[(size_t)(c + 1u) * nblocks] : n;
    const uint32_t nch = (e - s + kCellChunk - 1) / kCellChunk;
    uint32_t tot;
    const uint32_t c0 = block_excl_256<uint32_t>(nch, &tot);
    if (in) {
      b.cell_offsets[c] = s;
      b.cell_counts[c] = e - s;
      b.cell_nchunk[c] = nch;
      b.cell_chunk0[c] = c0;
    }
    return;
  }
  if (!keys || c >= b.total_cells) return;
  const uint32_t n = b.stats->cell_pairs;
  uint32_t s, e;
  {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (keys[mid] < c) lo = mid + 1; else hi = mid;
    }
    s = lo;
    hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (keys[mid] <= c) lo = mid + 1; else hi = mid;
    }
    e = lo;
  }
  b.cell_offsets[c] = s;
  b.cell_counts[c] = e - s;
  b.cell_nchunk[c] = (e - s + kCellChunk - 1) / kCellChunk;
}

// Expansion workgroups of the frame batch: sum over cells of ceil(count / kCellChunk).
__device__ __forceinline__ uint32_t cell_chunk_total(const Bufs& b) {
  const uint32_t last = b.total_cells - 1;
  return b.cell_chunk0[last] + b.cell_nchunk[last];
}


struct ChunkInfo {
  uint32_t cell, first, count, cx, cy;
  const FrameDesc* F;
};

// The cell and entry range of expansion chunk q (uniform per workgroup).
__device__ __forceinline__ ChunkInfo chunk_info(const Bufs& b, uint32_t q) {
  uint32_t lo = 0, hi = b.total_cells;  // last cell with chunk0 <= q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (b.cell_chunk0[mid] <= q) lo = mid; else hi = mid;
  }
  ChunkInfo ci;
  ci.cell = lo;
  const uint32_t k = q - b.cell_chunk0[lo];
  ci.first = b.cell_offsets[lo] + k * kCellChunk;
  ci.count = min(kCellChunk, b.cell_counts[lo] - k * kCellChunk);
  const uint32_t f = frame_of(b.cell_frame_start, b.num_frames, lo);
  ci.F = &b.frames[f];
  const uint32_t lc = lo - ci.F->cell_base;
  ci.cx = lc % ci.F->cells_x;
  ci.cy = lc / ci.F->cells_x;
  return ci;
}

// Bit (ly * 8 + lx) set for each tile (lx, ly) of cell (cx, cy) inside the tile box bd.
__device__ __forceinline__ unsigned long long cover_mask(uint32_t bd, uint32_t cx, uint32_t cy) {
  if (bd == HDSR_SENTINEL) return 0ull;
  uint32_t x0, y0, x1, y1;
  box_of(bd, &x0, &y0, &x1, &y1);
  const int bx = (int)(cx * kCellTiles), by = (int)(cy * kCellTiles);
  const int lx0 = max((int)x0 - bx, 0), lx1 = min((int)x1 - bx, 7);
  const int ly0 = max((int)y0 - by, 0), ly1 = min((int)y1 - by, 7);
  if (lx0 > lx1 || ly0 > ly1) return 0ull;
  const unsigned long long row = (unsigned long long)(((2u << lx1) - (1u << lx0)) & 0xFFu);
  unsigned long long m = 0ull;
#pragma unroll
  for (int r = 0; r < 8; ++r)
    if (r >= ly0 && r <= ly1) m |= row << (8 * r);
  return m;
}

// The rectangle [lx0, lx1] x [ly0, ly1] of cell (cx, cy)'s tiles (0..7 each) inside tile box bd;
// false when they do not meet.
__device__ __forceinline__ bool cell_rect(uint32_t bd, uint32_t cx, uint32_t cy, uint32_t* lx0, uint32_t* lx1,
                                          uint32_t* ly0, uint32_t* ly1) {
  if (bd == HDSR_SENTINEL) return false;
  uint32_t x0, y0, x1, y1;
  box_of(bd, &x0, &y0, &x1, &y1);
  const int bx = (int)(cx * kCellTiles), by = (int)(cy * kCellTiles);
  const int ax0 = max((int)x0 - bx, 0), ax1 = min((int)x1 - bx, 7);
  const int ay0 = max((int)y0 - by, 0), ay1 = min((int)y1 - by, 7);
  if (ax0 > ax1 || ay0 > ay1) return false;
  *lx0 = (uint32_t)ax0;
  *lx1 = (uint32_t)ax1;
  *ly0 = (uint32_t)ay0;
  *ly1 = (uint32_t)ay1;
  return true;
}

// Dyadic intervals over the 8 tiles of a cell row or column: node (L, i) covers [i 2^L, (i+1) 2^L),
// L = 0..3, stored at kDyOff[L] + i (15 nodes).  [a, b] splits into at most four disjoint nodes,
// returned as 4-bit node numbers packed low first; *n gets their count.
constexpr uint32_t kDyNodes = 15;
__device__ __forceinline__ uint32_t dyadic_pieces(uint32_t a, uint32_t b, uint32_t* n) {
  uint32_t packed = 0, k = 0, lo = a;
  const uint32_t hi = b + 1u;
  while (lo < hi) {
    uint32_t L = lo ? (uint32_t)__builtin_ctz(lo) : 3u;
    L = min(L, 3u);
    while ((1u << L) > hi - lo) --L;
    const uint32_t off = L == 0 ? 0u : L == 1 ? 8u : L == 2 ? 12u : 14u;
    packed |= (off + (lo >> L)) << (4u * k);
    ++k;
    lo += 1u << L;
  }
  *n = k;
  return packed;
}

// Counting walk: for the chunk's entries, the number covering each of the cell's 64 tiles, and each
// tile's min / max depth bits over them (the depth range that sets the tile's log frame).  Counts:
// every entry adds its tile rectangle to a 9 x 9 difference array (four LDS atomics), whose 2D
// prefix sums are the counts.  Depth range: every rectangle is split into dyadic x-intervals times
// dyadic y-intervals (at most 4 x 4 disjoint pieces; one for a rectangle spanning the cell) and the
// depth folded into each piece's min / max slot; a tile's min / max is then taken over the 16
// dyadic rectangles that contain it, exactly the pieces of the entries covering it.  Then one
// global atomic pair per tile and chunk.
__global__ __launch_bounds__(256) void k_cell_count(Bufs b, const uint32_t* cv) {
  __shared__ int32_t s_diff[9 * 9];
  __shared__ uint32_t s_dmin[kDyNodes * kDyNodes], s_dmax[kDyNodes * kDyNodes];
  const uint32_t q = blockIdx.x;
  const uint32_t nchunks = cell_chunk_total(b);
  if (q >= nchunks) return;  // uniform
  if (q == 0 && threadIdx.x == 0) b.stats->cell_chunks = nchunks;
  const ChunkInfo ci = chunk_info(b, q);
  const uint32_t t = threadIdx.x;
  if (t < 81) s_diff[t] = 0;
  if (t < kDyNodes * kDyNodes) {
    s_dmin[t] = 0xFFFFFFFFu;
    s_dmax[t] = 0u;
  }
  __syncthreads();
#pragma unroll 1
  for (uint32_t i = t; i < ci.count; i += 256) {
    const uint2 r = b.bdep[cv[ci.first + i]];  // bounds and depth bits in one 8-B record
    uint32_t lx0, lx1, ly0, ly1;
    if (!cell_rect(r.x, ci.cx, ci.cy, &lx0, &lx1, &ly0, &ly1)) continue;
    const uint32_t d = r.y;
    atomicAdd(&s_diff[ly0 * 9 + lx0], 1);
    atomicAdd(&s_diff[ly0 * 9 + lx1 + 1], -1);
    atomicAdd(&s_diff[(ly1 + 1) * 9 + lx0], -1);
    atomicAdd(&s_diff[(ly1 + 1) * 9 + lx1 + 1], 1);
    uint32_t nxp, nyp;
    const uint32_t xp = dyadic_pieces(lx0, lx1, &nxp), yp = dyadic_pieces(ly0, ly1, &nyp);
    for (uint32_t c = 0; c < nyp; ++c) {
      const uint32_t row = ((yp >> (4u * c)) & 15u) * kDyNodes;
      for (uint32_t a = 0; a < nxp; ++a) {
        const uint32_t slot = row + ((xp >> (4u * a)) & 15u);
        atomicMin(&s_dmin[slot], d);
        atomicMax(&s_dmax[slot], d);
      }
    }
  }
  __syncthreads();
  if (t < 9) {  // prefix along x of row t
    int32_t run = 0;
    for (uint32_t x = 0; x < 9; ++x) {
      run += s_diff[t * 9 + x];
      s_diff[t * 9 + x] = run;
    }
  }
  __syncthreads();
  if (t < 9) {  // then along y of column t
    int32_t run = 0;
    for (uint32_t y = 0; y < 9; ++y) {
      run += s_diff[y * 9 + t];
      s_diff[y * 9 + t] = run;
    }
  }
  __syncthreads();
  if (t < 64) {
    const uint32_t x = t & 7u, y = t >> 3;
    const uint32_t c = (uint32_t)s_diff[y * 9 + x];
    b.chunk_cnt[(size_t)q * 64 + t] = c;
    const FrameDesc& F = *ci.F;
    const uint32_t tx = ci.cx * kCellTiles + x, ty = ci.cy * kCellTiles + y;
    if (c && tx < F.tiles_x && ty < F.tiles_y) {
      uint32_t mn = 0xFFFFFFFFu, mx = 0u;
#pragma unroll
      for (uint32_t Ly = 0; Ly < 4; ++Ly) {
        const uint32_t row = ((Ly == 0 ? 0u : Ly == 1 ? 8u : Ly == 2 ? 12u : 14u) + (y >> Ly)) * kDyNodes;
#pragma unroll
        for (uint32_t Lx = 0; Lx < 4; ++Lx) {
          const uint32_t slot = row + (Lx == 0 ? 0u : Lx == 1 ? 8u : Lx == 2 ? 12u : 14u) + (x >> Lx);
          mn = min(mn, s_dmin[slot]);
          mx = max(mx, s_dmax[slot]);
        }
      }
      const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
      atomicMin(&b.tile_dmin[m], mn);
      atomicMax(&b.tile_dmax[m], mx);
    }
  }
}

// One 64-thread workgroup per cell: exclusive prefix of each tile's counts over the cell's chunks
// (in place), and the tile's total into tile_counts.
__global__ __launch_bounds__(64) void k_cell_prefix(Bufs b) {
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  const uint32_t f = frame_of(b.cell_frame_start, b.num_frames, c);
  const FrameDesc& F = b.frames[f];
  const uint32_t lc = c - F.cell_base;
  const uint32_t cx = lc % F.cells_x, cy = lc / F.cells_x;
  const uint32_t q0 = b.cell_chunk0[c], nq = b.cell_nchunk[c];
  uint32_t run = 0;
  // up to 32 chunks' counts are loaded before any of their prefixes is stored, so a cell of up
  // to 32k pairs takes one memory round trip
  for (uint32_t g = 0; g < nq; g += 32) {
    uint32_t v[32];
#pragma unroll
    for (uint32_t k = 0; k < 32; ++k) v[k] = g + k < nq ? b.chunk_cnt[(size_t)(q0 + g + k) * 64 + t] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < 32; ++k) {
      if (g + k < nq) b.chunk_cnt[(size_t)(q0 + g + k) * 64 + t] = run;
      run += v[k];
    }
  }
  const uint32_t tx = cx * kCellTiles + (t & 7u), ty = cy * kCellTiles + (t >> 3);
  if (tx < F.tiles_x && ty < F.tiles_y) b.tile_counts[F.morton_base + hdsr_morton2(tx, ty)] = run;
}

// Keeps every tile's range inside the pair capacity.  Only a frame whose references exceed the
// capacity is affected (its lists are cut; the host grows the capacity and renders it again),
// but every later stage then reads and writes inside the entry buffers.
//
// With the reference's caps, the large splats past the first 65,535 (unl: per tile) are counted
// for the offsets but never stored; they are the last entries of every tile's canonical list, so
// the tile keeps the first count - unl entries, and those below the cap.
__global__ void k_clamp_ranges(uint32_t* offsets, uint32_t* counts, uint32_t n, uint32_t cap, const uint32_t* unl) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = offsets[i], e = s + counts[i] - (unl ? unl[i] : 0u);
  const uint32_t cs = (uint32_t)(s < cap ? s : cap), ce = (uint32_t)(e < cap ? e : cap);
  offsets[i] = cs;
  counts[i] = ce - cs;
}

// The tile-count scan and the clamp above in one workgroup (n <= kOneGroupScan).  Thread t holds
// elements 16t .. 16t+15 in registers, all loaded at once (one memory round trip), and scans them
// in the thread, across its wave and across the 16 waves' sums; each element's range is clamped
// as it is stored.
__global__ __launch_bounds__(1024) void k_tile_ranges_one(uint32_t* __restrict__ counts, uint32_t* __restrict__ offsets,
                                                          uint32_t n, uint32_t cap, const uint32_t* __restrict__ unl) {
  __shared__ uint32_t wsum[16];
  const uint32_t t = threadIdx.x, lane = __lane_id(), w = t >> 6;
  const uint32_t i0 = t * 16u;
  // the thread's 16 elements are all inside: vector loads and stores (counts and offsets are
  // separate allocations, 16-byte aligned; the unlisted counts need not be)
  const bool whole = i0 + 16u <= n;
  uint32_t c[16], u[16];
  if (whole) {
#pragma unroll
    for (uint32_t k = 0; k < 16; k += 4) {
      const uint4 v = *reinterpret_cast<const uint4*>(counts + i0 + k);
      c[k] = v.x; c[k + 1] = v.y; c[k + 2] = v.z; c[k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) c[k] = i0 + k < n ? counts[i0 + k] : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) u[k] = (unl && i0 + k < n) ? unl[i0 + k] : 0u;
  uint32_t part = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) part += c[k];
  uint32_t inc = part;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {  // inclusive scan across the wave
    const uint32_t x = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += x;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t run = inc - part;  // exclusive offset of element 16t (u32, as the offsets are)
#pragma unroll
  for (uint32_t q = 0; q < 16; ++q)
    if (q < w) run += wsum[q];
  uint32_t o[16], r[16];
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    const uint64_t st = run, e = st + c[k] - u[k];
    const uint32_t cs = (uint32_t)(st < cap ? st : cap), ce = (uint32_t)(e < cap ? e : cap);
    o[k] = cs;
    r[k] = ce - cs;
    run += c[k];
  }
  if (whole) {
#pragma unroll
    for (uint32_t k = 0; k < 16; k += 4) {
      *reinterpret_cast<uint4*>(offsets + i0 + k) = make_uint4(o[k], o[k + 1], o[k + 2], o[k + 3]);
      *reinterpret_cast<uint4*>(counts + i0 + k) = make_uint4(r[k], r[k + 1], r[k + 2], r[k + 3]);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
      if (i0 + k < n) {
        offsets[i0 + k] = o[k];
        counts[i0 + k] = r[k];
      }
  }
}

// Writing walk, tile-major: the chunk's entries (splat, tile mask, ln depth) are staged in LDS;
// wave w then takes the cell's tiles 16w .. 16w+15 one at a time and walks all entries of the
// chunk in order, so each tile's run of this chunk is written as one contiguous range (whole
// cache lines instead of a few entries per store).
__global__ __launch_bounds__(256) void k_cell_write(Bufs b, const uint32_t* cv) {
  // one group of padding: the walk reads the group after the last (and discards it)
  __shared__ unsigned long long s_mask[kCellChunk + 64];
  __shared__ uint32_t s_g[kCellChunk + 64];
  __shared__ float s_ln[kCellChunk + 64];
  const uint32_t q = blockIdx.x;
  if (q >= cell_chunk_total(b)) return;  // uniform
  const ChunkInfo ci = chunk_info(b, q);
  const uint32_t t = threadIdx.x, w = t >> 6, lane = __lane_id();
  const FrameDesc& F = *ci.F;
  // lanes 0..15 of wave w fetch tile 16w + lane's write base and log-depth frame up front, so
  // the loads and the logarithms of all 16 tiles overlap the staging below
  uint32_t my_base = 0, my_lm = 0, my_sc = 0, my_ok = 0;
  if (lane < 16) {
    const uint32_t tt = w * 16 + lane;
    const uint32_t tx = ci.cx * kCellTiles + (tt & 7u), ty = ci.cy * kCellTiles + (tt >> 3);
    if (tx < F.tiles_x && ty < F.tiles_y) {
      const uint32_t m = F.morton_base + hdsr_morton2(tx, ty);
      my_base = b.tile_offsets[m] + b.chunk_cnt[(size_t)q * 64 + tt];
      float lm, sc;
      hdsr_log_frame(b.tile_dmin[m], b.tile_dmax[m], &lm, &sc);
      my_lm = hdsr_f2u(lm);
      my_sc = hdsr_f2u(sc * 8192.0f);  // the q20 scale
      my_ok = 1;
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < kCellChunk / 256; ++r) {
    const uint32_t i = r * 256 + t;
    unsigned long long mk = 0ull;
    uint32_t g = 0;
    float ln = 0.0f;
    if (i < ci.count) {
      g = cv[ci.first + i];
      const uint2 r = b.bln[g];  // bounds and ln(max(depth, 1e-6)), computed once per splat by the projection
      mk = cover_mask(r.x, ci.cx, ci.cy);
      ln = hdsr_u2f(r.y);
    }
    s_mask[i] = mk;
    s_g[i] = g;
    s_ln[i] = ln;
  }
  __syncthreads();
  const uint32_t cap = b.cap;
  const uint32_t ngroups = (ci.count + 63) / 64;
  // the tile's bit is tested in the one 32-bit half of the mask that holds it
  const uint32_t* s_mw = reinterpret_cast<const uint32_t*>(s_mask);
  for (uint32_t k = 0; k < 16; ++k) {
    const uint32_t tt = w * 16 + k;  // this wave's tile (uniform)
    if (!__builtin_amdgcn_readlane(my_ok, k)) continue;  // outside the tile grid
    uint32_t base = __builtin_amdgcn_readlane(my_base, k);
    const float lm = hdsr_u2f(__builtin_amdgcn_readlane(my_lm, k));
    const float sq = hdsr_u2f(__builtin_amdgcn_readlane(my_sc, k));
    const uint32_t half = tt >> 5, bit = 1u << (tt & 31u);
    // group gi's writes overlap the LDS reads of group gi + 1 (two register sets, no copies)
    auto put = [&](uint32_t mw, uint32_t g, float ln) {
      const bool in = (mw & bit) != 0u;  // entries past the count have an empty mask
      const unsigned long long mm = __ballot(in);
      if (in) {
        // base + covering lanes below this one
        const uint32_t dst = __builtin_amdgcn_mbcnt_hi((uint32_t)(mm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mm, base));
        if (dst < cap) {
          b.entries[dst] = g;
          b.qkey[dst] = hdsr_q20_scaled((ln - lm) * sq);  // the entry's key in this tile's frame
        }
      }
      base += (uint32_t)__popcll(mm);
    };
    uint32_t mwa = s_mw[2 * lane + half], ga = s_g[lane];
    float lna = s_ln[lane];
    for (uint32_t gi = 0; gi < ngroups; gi += 2) {
      const uint32_t ib = (gi + 1) * 64 + lane;  // at most the padding group
      const uint32_t mwb = s_mw[2 * ib + half], gb = s_g[ib];
      const float lnb = s_ln[ib];
      put(mwa, ga, lna);
      if (gi + 1 >= ngroups) break;  // uniform
      const uint32_t ia = (gi + 2) * 64 + lane;  // gi + 2 <= ngroups: at most the padding group
      mwa = s_mw[2 * ia + half];
      ga = s_g[ia];
      lna = s_ln[ia];
      put(mwb, gb, lnb);
    }
  }
}

// ---- the reference's WebGPU caps (HDSR_REF_CAPS, one frame) ---------------------------------

// Every large splat past the first 65,535 in rank order (the reference's list holds only those,
// gpu_classify_large_splats.wgsl:52,62) adds one to each tile of its box: those references are
// counted for the offsets but never stored.
__global__ __launch_bounds__(256) void k_unlisted(Bufs b) {
  const uint32_t nl = b.stats->large_splats;
  const uint32_t mb = b.frames[0].morton_base;
  for (uint32_t i = 65535u + blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
    uint32_t x0, y0, x1, y1;
    box_of(b.bounds[b.large_list[i]], &x0, &y0, &x1, &y1);
    for (uint32_t ty = y0; ty <= y1; ++ty)
      for (uint32_t tx = x0; tx <= x1; ++tx) atomicAdd(&b.tile_unl[mb + hdsr_morton2(tx, ty)], 1u);
  }
}

// A tile's stored entries are a prefix of its canonical list: its depth range (the log frame of
// its keys) is taken over them alone, and their keys are written again in that frame.
__global__ __launch_bounds__(256) void k_refcaps_frames(Bufs b) {
  __shared__ uint32_t s_mn[4], s_mx[4];
  const uint32_t m = blockIdx.x, t = threadIdx.x;
  const uint32_t s = b.tile_offsets[m], c = b.tile_counts[m];
  if (!c) return;  // uniform
  uint32_t mn = 0xFFFFFFFFu, mx = 0u;
  for (uint32_t i = t; i < c; i += 256) {
    const uint32_t d = b.bdep[b.entries[s + i]].y;
    mn = min(mn, d);
    mx = max(mx, d);
  }
  mn = wave_min(mn);
  mx = wave_max(mx);
  if ((t & 63u) == 0) {
    s_mn[t >> 6] = mn;
    s_mx[t >> 6] = mx;
  }
  __syncthreads();
  mn = min(min(s_mn[0], s_mn[1]), min(s_mn[2], s_mn[3]));
  mx = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
  if (t == 0) {
    b.tile_dmin[m] = mn;
    b.tile_dmax[m] = mx;
  }
  float lm, sc;
  hdsr_log_frame(mn, mx, &lm, &sc);
  for (uint32_t i = t; i < c; i += 256)
    b.qkey[s + i] = hdsr_q20_of((hdsr_u2f(b.bln[b.entries[s + i]].y) - lm) * sc);
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

// One tile's list entries and segments, at its class prefix p.
__device__ __forceinline__ void write_class(const Bufs& b, uint32_t t, uint32_t m, uint32_t s, uint32_t c, uint4 k,
                                            uint4 p) {
  if (k.x) {
    b.small_list[p.x] = m;
    b.segs[p.z] = Seg{s, s, c, 0xFFFFFFFFu};
  } else if (k.y) {
    b.large_list_t[p.y] = m;
    b.large_bases[p.y] = (b.ref_caps ? b.cap : b.stats->refs) + p.w;  // capacity + earlier large counts
    b.large_tiles[p.y] = LargeTile{m, s, c, p.w, p.z, 0u, t, 0u};
    b.tile_large[m] = p.y;
    for (uint32_t q = 0; q < k.z; ++q) {
      const uint32_t off = q * HDSR_MAX_TILE_ENTRIES;
      const uint32_t cnt = min(HDSR_MAX_TILE_ENTRIES, c - off);
      // with the caps, a chunk past the first 1,024 (p.z - p.x: the chunks of earlier large tiles)
      // keeps its bucket order (gpu_tile_bucket_sort.wgsl:126,138)
      const uint32_t keep = (b.ref_caps && p.z - p.x + q >= 1024u) ? kSegKeepOrder : 0u;
      b.segs[p.z + q] = Seg{p.w + off, s + off, cnt | keep, p.y};
    }
  }
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
  write_class(b, t, m, s, c, k, p);
}

// Classification in two launches (total_tiles <= kTwoPassTiles): every 256-tile workgroup sums its
// tiles' class vectors, then each workgroup adds the sums of the workgroups before it (at most 256,
// one per thread) to its own block scan and writes its tiles as k_classify_write does.
constexpr uint32_t kTwoPassTiles = 256u * 256u;
__global__ __launch_bounds__(256) void k_classify_part(Bufs b, uint4* part) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  uint4 k = make_uint4(0, 0, 0, 0);
  if (t < b.total_tiles) {
    uint32_t m, s, c;
    tile_info(b, t, &m, &s, &c);
    k = class_of(c);
  }
  uint4 tot;
  (void)block_excl_256<uint4>(k, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_classify_write2(Bufs b, const uint4* part) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const bool in = t < b.total_tiles;
  uint32_t m = 0, s = 0, c = 0;
  if (in) tile_info(b, t, &m, &s, &c);
  const uint4 before = threadIdx.x < blockIdx.x ? part[threadIdx.x] : make_uint4(0, 0, 0, 0);
  uint4 base, tot;
  (void)block_excl_256<uint4>(before, &base);  // the earlier workgroups' sum
  const uint4 k = in ? class_of(c) : make_uint4(0, 0, 0, 0);
  const uint4 p = base + block_excl_256<uint4>(k, &tot);
  if (!in) return;
  if (t == b.total_tiles - 1) {
    b.stats->small_tiles = p.x + k.x;
    b.stats->large_tiles = p.y + k.y;
    b.stats->segments = p.z + k.z;
    b.stats->chunks = p.z + k.z - (p.x + k.x);
    b.stats->scratch_used = p.w + k.w;
  }
  write_class(b, t, m, s, c, k, p);
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
  if (b.splat_blocks) return;  // the block scan clears them with the totals
  hipLaunchKernelGGL(k_clear_stats, dim3(1), dim3(64), 0, stream, b.stats);
}

void launch_scan_blocks(const Bufs& b, hipStream_t stream) {
  if (!b.splat_blocks) return;
  if (b.splat_blocks <= kOneGroupBlocks) {
    hipLaunchKernelGGL(k_scan_blocks_one, dim3(1), dim3(1024), 0, stream, b);
    return;
  }
  scan_u4_exclusive(b.block_cnt, b.block_pre, b.splat_blocks, b.scan_tmp, stream);
  scan_u4_exclusive(b.block_cc, b.block_ccpre, b.splat_blocks, b.scan_tmp, stream);
  hipLaunchKernelGGL(k_totals, dim3(1), dim3(1), 0, stream, b);
}

void launch_emit_pairs(const Bufs& b, hipStream_t stream) {
  if (!b.splat_blocks) return;
  hipLaunchKernelGGL(k_emit, dim3(b.splat_blocks), dim3(kSplatBlock), 0, stream, b);
}

// Passes and digit width of the cell-id radix sort; the sorted pairs end in the b buffers after an
// odd number of passes, in the a buffers after an even one.
static uint32_t cell_passes(const Bufs& b) {
  return b.cell_key_bits == 0 ? 1u : (b.cell_key_bits + kRadixBits - 1) / kRadixBits;
}
static const uint32_t* sorted_cell_keys(const Bufs& b) { return cell_passes(b) & 1u ? b.ckeys_b : b.ckeys_a; }
static const uint32_t* sorted_cell_vals(const Bufs& b) { return cell_passes(b) & 1u ? b.cvals_b : b.cvals_a; }

// The batch has (cell, splat) pairs to group: at least one splat block and one cell.
static bool has_pairs(const Bufs& b) { return b.splat_blocks && b.cell_cap && b.total_cells; }

void launch_group_pairs(const Bufs& b, hipStream_t stream) {
  if (!b.total_morton) return;
  const bool pairs = has_pairs(b);
  if (pairs)
    (void)radix_sort_pairs(b.ckeys_a, b.cvals_a, b.ckeys_b, b.cvals_b, &b.stats->cell_pairs, b.cell_cap,
                           b.cell_key_bits, b.radix_hist, b.radix_tmp, stream);
  // cell ranges and expansion chunks, per-chunk tile counts, per-cell prefixes -> tile counts
  const uint32_t cg = std::max((b.total_cells + 255) / 256, std::min((b.total_morton + 255) / 256, 1024u));
  // one radix pass: the cells are its digits, and its scanned histogram holds their ranges
  const bool one_pass = pairs && cell_passes(b) == 1 && b.total_cells <= kRadixBuckets;
  const uint32_t nblocks = (b.cell_cap + kRadixBlockItems - 1) / kRadixBlockItems;
  hipLaunchKernelGGL(k_cell_ranges, dim3(cg), dim3(256), 0, stream, b, pairs ? sorted_cell_keys(b) : nullptr,
                     one_pass ? (const uint32_t*)b.radix_hist : nullptr, nblocks);
  if (!pairs) return;  // every tile count is zero: the later stages find no entries
  if (!one_pass) scan_u32_exclusive(b.cell_nchunk, b.cell_chunk0, b.total_cells, b.radix_tmp, stream);
  hipLaunchKernelGGL(k_cell_count, dim3(b.max_chunks), dim3(256), 0, stream, b, sorted_cell_vals(b));
  hipLaunchKernelGGL(k_cell_prefix, dim3(b.total_cells), dim3(64), 0, stream, b);
  if (b.ref_caps && b.tile_unl)
    hipLaunchKernelGGL(k_unlisted, dim3(1024), dim3(256), 0, stream, b);
}

void launch_tile_ranges(const Bufs& b, hipStream_t stream) {
  if (!b.total_morton) return;
  if (b.total_morton <= kOneGroupScan) {
    hipLaunchKernelGGL(k_tile_ranges_one, dim3(1), dim3(1024), 0, stream, b.tile_counts, b.tile_offsets,
                       b.total_morton, b.cap, (const uint32_t*)b.tile_unl);
  } else {
    scan_u32_exclusive(b.tile_counts, b.tile_offsets, b.total_morton, b.radix_tmp, stream);
    hipLaunchKernelGGL(k_clamp_ranges, dim3((b.total_morton + 255) / 256), dim3(256), 0, stream, b.tile_offsets,
                       b.tile_counts, b.total_morton, b.cap, (const uint32_t*)b.tile_unl);
  }
  if (has_pairs(b))
    hipLaunchKernelGGL(k_cell_write, dim3(b.max_chunks), dim3(256), 0, stream, b, sorted_cell_vals(b));
  if (b.ref_caps && has_pairs(b))  // the stored lists' own depth ranges and keys
    hipLaunchKernelGGL(k_refcaps_frames, dim3(b.total_morton), dim3(256), 0, stream, b);
}

size_t radix_hist_words(uint32_t cap) {
  return (size_t)((cap + kRadixBlockItems - 1) / kRadixBlockItems) * kRadixBuckets;
}

bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, const uint32_t* count,
                      uint32_t cap, uint32_t key_bits, uint32_t* hist, uint32_t* hist_tmp, hipStream_t stream) {
  const uint32_t nblocks = (cap + kRadixBlockItems - 1) / kRadixBlockItems;
  if (!nblocks) return false;
  const uint32_t passes = key_bits == 0 ? 1u : (key_bits + kRadixBits - 1) / kRadixBits;
  const uint32_t bits = key_bits == 0 ? 1u : (key_bits + passes - 1) / passes;
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t *ks = keys, *vs = vals, *kd = keys_tmp, *vd = vals_tmp;
  for (uint32_t p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(256), 0, stream, ks, hist, count, nblocks, p * bits, mask);
    scan_u32_exclusive(hist, hist, nblocks * kRadixBuckets, hist_tmp, stream);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(256), 0, stream, ks, vs, kd, vd, hist, count, nblocks,
                       p * bits, bits);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  return (passes & 1u) != 0;
}

void launch_classify(const Bufs& b, hipStream_t stream) {
  if (!b.total_tiles) return;
  // the class vector and its prefix live in the uint4 scan area after the block prefix scratch
  uint4* cls = b.scan_tmp + scan_u4_tmp_words(b.splat_blocks);
  uint4* pre = cls + b.total_tiles;
  uint4* tmp = pre + b.total_tiles;
  const uint32_t grid = (b.total_tiles + 255) / 256;
  if (b.total_tiles <= kTwoPassTiles) {  // the workgroup sums take the class vector's place
    hipLaunchKernelGGL(k_classify_part, dim3(grid), dim3(256), 0, stream, b, cls);
    hipLaunchKernelGGL(k_classify_write2, dim3(grid), dim3(256), 0, stream, b, (const uint4*)cls);
    return;
  }
  hipLaunchKernelGGL(k_classify_count, dim3(grid), dim3(256), 0, stream, b, cls);
  scan_u4_exclusive(cls, pre, b.total_tiles, tmp, stream);
  hipLaunchKernelGGL(k_classify_write, dim3(grid), dim3(256), 0, stream, b, (const uint4*)pre);
}

}  // namespace hdsr
