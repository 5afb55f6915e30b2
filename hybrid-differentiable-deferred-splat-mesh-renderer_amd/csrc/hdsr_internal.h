// hdsr_internal.h - device data layout and kernel launchers of the splat compositor.
//
// One launch of each stage processes a batch of V frames (V = 1 for the plain drop-in; up to
// 64 in the serving form).  Every per-splat, per-pair and per-tile array is one flat array over
// the batch; a frame owns a contiguous range of it:
//   splats  g in [splat_base, splat_base + n)          (splat_base a multiple of 256)
//   Morton  m in [morton_base, morton_base + morton_count)
//   tiles   t in [tile_base, tile_base + tiles_x * tiles_y)   (row-major inside the frame)
//   cells   c in [cell_base, cell_base + cells_x * cells_y) (8 x 8-tile cells, row-major)
// Tile lists are built in two levels: every visible splat emits one (cell, rank) pair per cell
// its tile box touches, in canonical order (small splats by rank, then large splats by rank); one
// stable radix sort on the global cell id groups them; each cell's list is then expanded into
// its 64 tiles' lists with ballots, which keeps the canonical order inside every tile.  The
// entry list is laid out by global Morton slot (morton_base + morton(tx, ty)).
//
// Projected splat records (the reference's 14-word cache, gpu_project_splats.wgsl:2-16, split
// by who reads it):
//   hot[g]  32 B  {px, py, k0, k1 | k2, depth, f16 rg, f16 b | f16 opacity}  - composite reads
//   cold[g] 32 B  {w5 material, w6 radius|opacity, w8 normal, w13 conf | w11, w12, 0, 0}
//   bounds[g]     w7 tile box (u8 x 4), HDSR_SENTINEL if invisible           - count / emit
//   depth[g]      depthNdc bits                                              - per-tile keys
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hdsr_fmath.h"

namespace hdsr {

constexpr uint32_t kSplatBlock = 256;        // splats per project/emit workgroup
constexpr uint32_t kRadixBits = 8;           // digit width of one radix pass
constexpr uint32_t kRadixBuckets = 1u << kRadixBits;
constexpr uint32_t kRadixBlockItems = 4096;  // pairs per radix workgroup (256 threads x 16)
constexpr uint32_t kScanChunk = 4096;        // elements per workgroup of the generic scan
constexpr uint32_t kMaxFrames = 64;
constexpr uint32_t kCellTiles = 8;            // a cell is 8 x 8 tiles
constexpr uint32_t kCellChunk = 1024;         // cell-list entries per expansion workgroup
constexpr uint32_t kMaxCellsPerSplat = 9;     // a tile box spans at most 17 tiles per axis
// A deferred frame's resumed tiles: entries before this list index get their quadrant masks from
// k_resume_masks (a multiple of 4096 past the first chunk); the rest are tested by the rasterizer.
#define HDSR_RESUME_MASKED (4u * HDSR_MAX_TILE_ENTRIES)

// Floats per splat of a shData row at SH degree deg: dc (3) + 3K coefficients + emissive (3),
// K = (deg+1)^2 - 1 (gpu_project_splats.wgsl:294-296).
inline uint32_t sh_row_floats(uint32_t deg) { return 3u + 3u * ((deg + 1u) * (deg + 1u) - 1u) + 3u; }

// The SH degree of a bind: deg_in 0..3, or 0xFFFFFFFF (take it from sh_len, the floats in shData,
// which must be count * sh_row_floats(d) for exactly one d).  A stated sh_len must hold the rows.
// Returns false with *why set when the buffers cannot be read at that degree.
inline bool resolve_sh_degree(uint32_t count, uint32_t deg_in, uint64_t sh_len, uint32_t* deg_out,
                              const char** why) {
  if (deg_in == 0xFFFFFFFFu) {
    if (count == 0) {
      *deg_out = 0;
      return true;
    }
    for (uint32_t d = 0; d <= 3; ++d)
      if (sh_len == (uint64_t)count * sh_row_floats(d)) {
        *deg_out = d;
        return true;
      }
    *why = "shData length matches no SH degree: it must hold count * (6 + 3K) floats, K = (d+1)^2 - 1";
    return false;
  }
  if (deg_in > 3) {
    *why = "sh_degree must be 0..3";
    return false;
  }
  if (sh_len != 0 && sh_len < (uint64_t)count * sh_row_floats(deg_in)) {
    *why = "shData is shorter than count * (6 + 3K) floats for the stated sh_degree";
    return false;
  }
  *deg_out = deg_in;
  return true;
}

struct FrameDesc {
  // scene (device pointers)
  const float* pos;
  const float* logscale;
  const float* rot;
  const float* opacity;
  const uint32_t* sorted;
  const uint32_t* material;
  const float* normals;
  const float* shdata;
  uint32_t normals_len;
  uint32_t n;
  uint32_t sh_degree;
  uint32_t sh_stride;  // floats per splat in shdata: 3 + 3K + 3
  // The same scene as one 128-byte-aligned record per splat (a bind-time copy of a scene the
  // library owns), or null; see packed_stride4.
  const float4* packed;
  uint32_t pk_stride4;  // float4 slots per record
  // camera (parsed from the 256-byte uniform block)
  float vp[16];
  float view[16];
  float nm[9];
  float cam[3];
  float vw, vh, tile, splat_scale, fx, fy;
  uint32_t tiles_x, tiles_y;
  // batch layout
  uint32_t splat_base, morton_base, morton_count, tile_base;
  uint32_t cells_x, cells_y, cell_base;
  uint32_t width, height;
  // outputs (device pointers)
  float* rgba;           // width*height*4
  float* out_depth;      // width*height
  uint32_t* gb_normal;   // width*height, or null
  uint32_t* gb_material; // width*height*4, or null
  uint16_t* rgba16;      // width*height*4 binary16, or null
  // differentiable mode: opacity holds logits, the rasterizer reads the fp32 record (drec) and
  // keeps each pixel's final transmittance and one-past-last contributing entry for the backward
  uint32_t diff;
  float* t_final;        // width*height, or null
  uint32_t* last_idx;    // width*height, or null
};

// Counters of one batch, written by the kernels (device) and copied to pinned host memory.
struct DevStats {
  uint32_t small_pairs;   // pairs from splats covering <= 16 tiles
  uint32_t large_pairs;   // pairs from splats covering > 16 tiles
  uint32_t large_splats;
  uint32_t visible;
  uint32_t refs;          // small_pairs + large_pairs
  uint32_t overflow;      // 1 if refs exceeded the pair capacity (frame must be re-run)
  uint32_t small_tiles;
  uint32_t large_tiles;
  uint32_t segments;      // small tiles with >= 1 entry + chunks of large tiles
  uint32_t chunks;
  uint32_t scratch_used;
  uint32_t composite_refs;      // entries the rasterizer loaded (up to each tile's saturation)
  uint32_t composite_max_refs;  // largest per-tile count of those
  uint32_t cell_pairs;          // (cell, rank) pairs
  uint32_t cell_chunks;         // expansion workgroups needed (sum over cells of ceil(count / kCellChunk))
  uint32_t resumed_refs;        // deferred frames: entries after the first chunk of the resumed tiles
  uint32_t pend_tiles;          // deferred frames: tiles the first round left unsaturated (pend_tiles[])
  uint32_t pend_segs;           // deferred frames: their chunks after the first (pend_segs[])
  uint32_t resumed_max;         // deferred frames: entries of the largest resumed tile
};

struct Seg {
  uint32_t src;    // first entry, in entries[] (small tile) or scratch[] (chunk)
  uint32_t dst;    // first entry in entries[]
  uint32_t count;  // <= 4096, | kSegKeepOrder for a chunk copied unsorted
  uint32_t large;  // index into the large-tile table, or 0xFFFFFFFF for a small tile
};
// Seg.count flag: with the reference's caps, a chunk past the first 1,024 keeps its bucket order.
constexpr uint32_t kSegKeepOrder = 0x80000000u;

struct LargeTile {
  uint32_t morton;   // global Morton slot
  uint32_t start;    // first entry in entries[]
  uint32_t count;
  uint32_t scratch;  // first slot in scratch[]
  uint32_t seg;      // the segment of its first chunk (the others follow it)
  uint32_t pending;  // deferred frames: 1 once the first round left the tile unsaturated (resumed in round 2)
  uint32_t tile;     // global tile index
  uint32_t pad3;
};

// Device buffers of one batch (all owned by the host context).
struct Bufs {
  const FrameDesc* frames;   // [V]
  const uint32_t* block_frame_start;  // [V+1] first splat block of each frame (host-built)
  const uint32_t* tile_frame_start;   // [V+1] first tile of each frame
  const uint32_t* cell_frame_start;   // [V+1] first cell of each frame
  uint32_t num_frames;
  uint32_t splat_blocks;     // sum of padded splats / 256
  uint32_t total_tiles;      // sum of tiles
  uint32_t total_morton;     // sum of Morton spans
  uint32_t total_cells;      // sum of cells
  uint32_t cell_key_bits;    // bits needed for a global cell id
  uint4* hot;                // [splat_blocks*256*2]
  uint4* cold;               // [splat_blocks*256*2]
  float4* drec;              // [splat_blocks*256*3] fp32 record of the differentiable mode, or null:
                             //   {px, py, k0, k1 | k2, opacity, r, g | b, depth, 0, 0}
  float* gacc;               // [splat_blocks*256*9] dL/d(px, py, k0, k1, k2, opacity, r, g, b), or null
  uint32_t* bounds;          // [splat_blocks*256]
  uint32_t* depth;           // [splat_blocks*256]
  uint2* bdep;               // [splat_blocks*256] (bounds, depth bits): the counting walk's gather per pair
  uint2* bln;                // [splat_blocks*256] (bounds, ln(max(depth, 1e-6)) bits): the list write's gather
  uint4* block_cnt;          // [splat_blocks] per block (small pairs, large pairs, large splats, visible)
  uint4* block_pre;          // [splat_blocks] exclusive prefix of block_cnt
  uint4* block_cc;           // [splat_blocks] per block (small-splat cell pairs, large-splat cell pairs, 0, 0)
  uint4* block_ccpre;        // [splat_blocks] exclusive prefix of block_cc
  uint4* scan_tmp;           // scratch of the uint4 scan
  uint32_t* large_list;      // [splat_blocks*256] sort ranks (global g) of large splats
  uint32_t* ckeys_a;         // [cell_cap] cell pair keys (global cell id), emission order
  uint32_t* cvals_a;         // [cell_cap] cell pair values (global splat index g)
  uint32_t* ckeys_b;         // [cell_cap] radix ping-pong
  uint32_t* cvals_b;         // [cell_cap]
  uint32_t cell_cap;         // kMaxCellsPerSplat per splat slot
  uint32_t* cell_offsets;    // [total_cells] first entry of each cell in the grouped cell list
  uint32_t* cell_counts;     // [total_cells]
  uint32_t* cell_nchunk;     // [total_cells] expansion chunks of each cell
  uint32_t* cell_chunk0;     // [total_cells] first expansion chunk of each cell (scan of cell_nchunk)
  uint32_t* chunk_cnt;       // [max_chunks * 64] per chunk and tile of its cell: count, then prefix
  uint32_t max_chunks;
  uint32_t* entries;         // [cap] final per-tile order (global g)
  uint32_t* keys20;          // [cap] (q20 << 12) | local index, or null
  uint32_t* entries_canon;   // [cap] entries before the per-tile sort, or null
  uint32_t* scratch;         // [cap] bucket-ordered copies of large tiles
  uint32_t* qkey;            // [cap] q20 key of each entry in its tile's log-depth frame (same order as entries);
                             //   in a deferred frame's round 2, a resumed tile's later chunks hold each
                             //   sorted entry's reach_quadrants mask instead (the keys are dead by then)
  uint32_t* qkey_s;          // [cap] q20 key of each scratch entry
  uint32_t* tile_dmin;       // [total_morton] min / max depth bits over each tile's entries (u32 order,
  uint32_t* tile_dmax;       //   as the reference's atomicMin / atomicMax)
  uint32_t cap;              // pair capacity
  uint32_t* radix_hist;      // [256 * radix blocks] digit-major (radix sort of the cell pairs)
  uint32_t* radix_tmp;       // scratch of the u32 scans (radix histogram, cell chunks, tile counts)
  uint32_t* tile_offsets;    // [total_morton]
  uint32_t* tile_counts;     // [total_morton]
  uint32_t* small_list;      // [total_tiles] global Morton slots
  uint32_t* large_list_t;    // [total_tiles] global Morton slots
  uint32_t* large_bases;     // [total_tiles]
  LargeTile* large_tiles;    // [total_tiles]
  Seg* segs;                 // [seg capacity]
  uint32_t seg_cap;
  DevStats* stats;
  uint32_t packed;           // every frame reads its scene from packed records (host-set, selects the kernel)
  uint32_t bucket_wide;      // large tiles average > 16k entries: one 1024-thread workgroup per tile (host-set)
  uint32_t first_chunks;     // deferred frames: round 1 selects the large tiles' first chunks without the
                             //   bucket scatter (k_first_chunks); rounds 2 and 3 then scatter (host-set)
  // Deferred chunk sort (host-set round; 0 = the whole frame in one pass).  Round 1 sorts the small
  // tiles and the first chunk of every large tile, and rasterizes each large tile up to the end of
  // that chunk; a tile left unsaturated there saves its pixel state and appends itself to
  // pend_tiles[] and its other chunks to pend_segs[].  Round 2 sorts those chunks and resumes
  // those tiles, walking the two lists only.  Round 3 sorts every chunk
  // after the first: it completes the lists of a deferred frame for a reader of the whole lists.
  // In round 1 a large tile's first chunk is selected and sorted from the tile's bucket
  // histogram into first[] (k_first_chunks), so its canonical list stays in entries[] until the
  // bucket scatter of round 2 (the resumed tiles) or round 3 (the others) reads it.
  uint32_t round;
  uint32_t ref_caps;          // the reference's WebGPU caps enforced (HDSR_REF_CAPS; one frame)
  uint32_t* tile_unl;         // [total_morton] with ref_caps: large splats past the first 65,535 per tile
  uint32_t* tile_large;      // [total_morton] large tiles only: index into large_tiles[]
  uint32_t* pend_tiles;      // [total_tiles] deferred frames: global tile indices resumed in round 2
  uint32_t* pend_segs;       // [seg capacity] deferred frames: the segments round 2 sorts
  uint32_t* pend_large;      // [total_tiles] deferred frames: large-tile index of each pend_tiles[] entry
  uint32_t* pend_aux;        // [2 * total_tiles] deferred frames: round 2's loaded entries and finished quadrants per pend_tiles[] entry
  uint32_t* first;           // [first_tiles * 4096] deferred frames: large tile l's sorted first chunk at l * 4096
  uint32_t* large_order;     // [total_tiles] large-tile indices by decreasing size (k_first_chunks' order)
  uint32_t first_tiles;      // large tiles first[] holds
  float* px_state;           // [total_tiles*256*6] T, r, g, b, depth weight, weight sum per pixel (planar)
  // The projection writes cold[] (read only by the G-buffer rasterizer and the 14-word debug
  // readback) only when set.
  uint32_t write_cold;
};

// Packed splat record, in floats: [0..3] rotation wxyz, [4..6] position, [7..9] log-scale,
// [10] opacity, [11] material bits, [12, 12 + sh_stride) the shData row, zero padding to a
// multiple of 32 floats.  SH3 (51 floats) fills 256 bytes, SH0 128: a splat gathered through
// sortedIndices costs two (one) 128-byte lines instead of one line per attribute array.
inline uint32_t packed_stride4(uint32_t sh_stride) { return ((12u + sh_stride + 31u) / 32u) * 8u; }
// Deferred consumers of the G-buffer (k_deferred.hip).
void launch_screen_space_normals(const float* depth, const uint32_t* src, uint32_t* out, uint32_t w, uint32_t h,
                                 const float* inv16, const float* cam3, uint32_t mode, float g0, float g1,
                                 hipStream_t stream);
// Returns -1 (nothing launched) for a radius outside [0, 8].
int launch_bilateral_normals(const uint32_t* in, const float* depth, uint32_t* out, uint32_t w, uint32_t h,
                             float spatial_sigma, float depth_sigma, float normal_sigma, int radius,
                             hipStream_t stream);
// The deferred lighting pass's Params block (src/splatRenderer.ts:505-523), 44 floats (k_lighting.hip).
struct LightingBlock {
  float f[44];
};
void launch_deferred_lighting(const LightingBlock& p, const float* rgba, const float* depth, const uint32_t* normal,
                              const uint32_t* mat, const float* ao, const float* bloom, const float* lut,
                              uint16_t* out, uint32_t w, uint32_t h, hipStream_t stream);
void launch_pack_scene(const float* pos, const float* logscale, const float* rot, const float* opacity,
                       const uint32_t* material, const float* shdata, uint32_t n, uint32_t sh_stride, float4* out,
                       hipStream_t stream);

// Launchers (one per pipeline stage); every one enqueues on 'stream' and returns.
void launch_clear_stats(const Bufs& b, hipStream_t stream);
void launch_project(const Bufs& b, hipStream_t stream);
void launch_scan_blocks(const Bufs& b, hipStream_t stream);
void launch_emit_pairs(const Bufs& b, hipStream_t stream);    // (cell, rank) pairs
void launch_group_pairs(const Bufs& b, hipStream_t stream);   // radix on the cell id, tile counts
void launch_tile_ranges(const Bufs& b, hipStream_t stream);   // tile offsets, expansion into tile lists
void launch_classify(const Bufs& b, hipStream_t stream);
void launch_bucket(const Bufs& b, hipStream_t stream);        // rounds 0, 2, 3: the whole bucket order
void launch_first_chunks(const Bufs& b, hipStream_t stream);  // round 1: large tiles' first chunks + small tiles
void launch_segment_sort(const Bufs& b, hipStream_t stream);
void launch_composite(const Bufs& b, hipStream_t stream, bool gbuffer, bool f16);
void launch_reset_pending(const Bufs& b, hipStream_t stream);  // empties the round-2 lists (re-run round 1)
// Differentiable mode (one frame, the batch's frame 0, with n splats): backward of the
// rasterizer into gacc, then of the projection into the parameter gradients, which are
// accumulated (added to) and indexed by splat id.
struct DiffGrads {
  float* means;           // 3N
  float* log_scales;      // 3N
  float* rotations;       // 4N
  float* opacity_logits;  // N
  float* sh_data;         // N * sh_stride (the emissive tail gets no gradient)
};
void launch_backward(const Bufs& b, uint32_t n, const float* grad_rgba, const DiffGrads& out, hipStream_t stream);
// Rebuilds the reference's 14-word projection cache of frame f into dst (debug readback).
void launch_cache14(const Bufs& b, uint32_t frame, uint32_t n, uint32_t* dst, hipStream_t stream);

// Stable LSD radix sort of (key, value) pairs in place of keys / vals (ping-pong through the tmp
// buffers), *count pairs (device) of at most cap, keys below 2^key_bits, digits of at most 8 bits.
// hist needs radix_hist_words(cap) words, hist_tmp scan_u32_tmp_words of that.  Returns true when
// the sorted pairs end in the tmp buffers (an odd number of passes).
size_t radix_hist_words(uint32_t cap);
bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, const uint32_t* count,
                      uint32_t cap, uint32_t key_bits, uint32_t* hist, uint32_t* hist_tmp, hipStream_t stream);
// Back-to-front order of n splats for one view (see k_depthsort.hip): out[rank] = splat id.
// work holds depth_sort_padded(n) (key, index) pairs.
uint32_t depth_sort_padded(uint32_t n);
void launch_depth_sort(const float* positions, uint32_t n, const float* view16, uint32_t* out, uint2* work,
                       hipStream_t stream);

// Generic exclusive scans used by several stages (hierarchical: reduce, scan of partials, downsweep).
size_t scan_u32_tmp_words(uint32_t n);
void scan_u32_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t stream);
size_t scan_u4_tmp_words(uint32_t n);
void scan_u4_exclusive(const uint4* in, uint4* out, uint32_t n, uint4* tmp, hipStream_t stream);

// Wave64 helpers ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
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

// Which frame owns item 'x' given the ascending starts table [V+1] (uniform per workgroup).
__device__ __forceinline__ uint32_t frame_of(const uint32_t* starts, uint32_t nframes, uint32_t x) {
  uint32_t lo = 0, hi = nframes;  // starts[lo] <= x < starts[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace hdsr

#define HIP_CHECK_LAUNCH() (void)hipGetLastError()
