// engine.hip - batch layout, buffer growth and the stage sequence of the compositor.
//
// Stage order per batch (reference pass order: gpuTileSplatCompositor.ts:1023-1179):
//   clear stats -> project (+count, +large classify) -> scan of block counts -> emit pairs
//   -> radix grouping on the Morton key -> tile ranges -> classify -> bucket cascade
//   -> segment sort (small tiles + chunks) -> composite
// The reference's buffer clears and indirect-argument passes have no counterpart: counts come
// from scans, and launches that depend on device-side totals use grid-stride loops.
#include "engine.h"

#include <algorithm>
#include <cstring>

namespace hdsr {

static inline uint32_t align_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

void parse_uniforms(const void* p, FrameDesc* F) {
  const float* f = static_cast<const float*>(p);
  const uint32_t* u = static_cast<const uint32_t*>(p);
  memcpy(F->vp, f, 64);
  F->vw = f[16];
  F->vh = f[17];
  F->tile = f[18];
  F->tiles_x = u[20];
  F->tiles_y = u[21];
  F->splat_scale = f[24];
  F->cam[0] = f[28];
  F->cam[1] = f[29];
  F->cam[2] = f[30];
  memcpy(F->view, f + 32, 64);
  F->fx = f[48];
  F->fy = f[49];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) F->nm[c * 3 + r] = f[52 + c * 4 + r];
}

Layout Engine::layout(FrameDesc* frames, uint32_t V) {
  Layout L;
  L.num_frames = V;
  L.block_start.resize(V + 1);
  L.tile_start.resize(V + 1);
  L.cell_start.resize(V + 1);
  uint32_t sb = 0, mb = 0, tb = 0, cb = 0;
  L.packed = V > 0;
  for (uint32_t v = 0; v < V; ++v) {
    FrameDesc& F = frames[v];
    L.packed &= F.packed != nullptr || F.n == 0;  // an empty frame reads no splat
    F.splat_base = sb;
    F.morton_base = mb;
    F.tile_base = tb;
    F.morton_count = (F.tiles_x && F.tiles_y) ? hdsr_morton2(F.tiles_x - 1, F.tiles_y - 1) + 1 : 0;
    F.cells_x = (F.tiles_x + kCellTiles - 1) / kCellTiles;
    F.cells_y = (F.tiles_y + kCellTiles - 1) / kCellTiles;
    F.cell_base = cb;
    L.block_start[v] = sb / kSplatBlock;
    L.tile_start[v] = tb;
    L.cell_start[v] = cb;
    sb += align_up(std::max(F.n, 1u), kSplatBlock);
    mb += F.morton_count;
    tb += F.tiles_x * F.tiles_y;
    cb += F.cells_x * F.cells_y;
  }
  L.block_start[V] = sb / kSplatBlock;
  L.tile_start[V] = tb;
  L.cell_start[V] = cb;
  L.splat_blocks = sb / kSplatBlock;
  L.total_tiles = tb;
  L.total_morton = mb;
  L.total_cells = cb;
  uint32_t bits = 0;
  while (bits < 32 && (1ull << bits) < (uint64_t)cb) ++bits;
  L.cell_key_bits = bits;
  return L;
}

Engine::~Engine() { release(); }

void Engine::release() {
  drop_graphs();
  if (gstage_) (void)hipHostFree(gstage_);
  gstage_ = nullptr;
  gstage_bytes_ = 0;
  void* ptrs[] = {b_.hot, b_.cold, b_.drec, b_.gacc, b_.bounds, b_.depth, b_.bdep, b_.bln, b_.block_cnt, b_.block_pre,
                  b_.block_cc, b_.block_ccpre, b_.scan_tmp, b_.large_list, b_.ckeys_a, b_.cvals_a, b_.ckeys_b,
                  b_.cvals_b, b_.cell_offsets, b_.chunk_cnt, b_.entries, b_.keys20, b_.entries_canon,
                  b_.scratch, b_.qkey, b_.qkey_s, b_.tile_dmin,
                  b_.radix_hist, b_.radix_tmp, b_.tile_offsets, b_.tile_counts, b_.small_list, b_.large_list_t,
                  b_.large_bases, b_.large_tiles, b_.segs, b_.stats, b_.tile_large, b_.pend_tiles, b_.pend_segs, b_.px_state,
                  b_.pend_large, b_.first, b_.large_order, b_.pend_aux,
                  d_frames_};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b_ = Bufs{};
  d_frames_ = nullptr;
  d_starts_ = nullptr;
  for (int i = 0; i < kRing; ++i) {
    if (stage_[i]) (void)hipHostFree(stage_[i]);
    if (stage_ev_[i]) (void)hipEventDestroy(stage_ev_[i]);
    stage_[i] = nullptr;
    stage_ev_[i] = nullptr;
  }
  if (hstats_) (void)hipHostFree(hstats_);
  hstats_ = nullptr;
  if (done_ev_) (void)hipEventDestroy(done_ev_);
  done_ev_ = nullptr;

  for (hipEvent_t e : prof_ev_) (void)hipEventDestroy(e);
  prof_ev_.clear();
  prof_head_ = prof_count_ = 0;
  n_splat_ = n_blocks_ = n_scan4_ = n_cap_ = n_hist_ = n_hist_tmp_ = 0;
  n_morton_ = n_tiles_ = n_seg_ = n_frames_ = n_debug_ = n_diff_ = 0;
  n_cellpairs_ = n_cells_ = n_chunkcnt_ = 0;
  n_first_ = 0;
}

template <typename T>
int Engine::grow(T** p, size_t* have, size_t need) {
  if (need <= *have && *p) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  ++epoch_;
  const size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
  if (hipMalloc((void**)p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    err = "device allocation of " + std::to_string(bytes) + " bytes failed";
    return HDSR_E_OOM;
  }
  *have = need;
  return 0;
}

int Engine::ensure(const Layout& L, bool debug) {
  int rc = 0;
  const size_t ns = (size_t)L.splat_blocks * kSplatBlock;
  size_t have;
  // per-splat arrays (sized together)
  if (ns > n_splat_ || !b_.hot) {
    have = n_splat_;
    if ((rc = grow(&b_.hot, &have, ns * 2))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.cold, &have, ns * 2))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.bounds, &have, ns))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.depth, &have, ns))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.large_list, &have, ns))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.bdep, &have, ns))) return rc;
    have = n_splat_;
    if ((rc = grow(&b_.bln, &have, ns))) return rc;
    n_splat_ = ns;
  }
  if (diff_ && (ns > n_diff_ || !b_.drec)) {
    have = n_diff_;
    if ((rc = grow(&b_.drec, &have, ns * 3))) return rc;
    have = n_diff_;
    if ((rc = grow(&b_.gacc, &have, ns * 9))) return rc;
    n_diff_ = ns;
  }
  if (L.splat_blocks > n_blocks_ || !b_.block_cnt) {
    have = n_blocks_;
    if ((rc = grow(&b_.block_cnt, &have, L.splat_blocks))) return rc;
    have = n_blocks_;
    if ((rc = grow(&b_.block_pre, &have, L.splat_blocks))) return rc;
    have = n_blocks_;
    if ((rc = grow(&b_.block_cc, &have, L.splat_blocks))) return rc;
    have = n_blocks_;
    if ((rc = grow(&b_.block_ccpre, &have, L.splat_blocks))) return rc;
    n_blocks_ = L.splat_blocks;
  }
  // cell pairs: at most kMaxCellsPerSplat per splat slot
  const size_t ncp = ns * kMaxCellsPerSplat;
  if (ncp > n_cellpairs_ || !b_.ckeys_a) {
    have = n_cellpairs_;
    if ((rc = grow(&b_.ckeys_a, &have, ncp))) return rc;
    have = n_cellpairs_;
    if ((rc = grow(&b_.cvals_a, &have, ncp))) return rc;
    have = n_cellpairs_;
    if ((rc = grow(&b_.ckeys_b, &have, ncp))) return rc;
    have = n_cellpairs_;
    if ((rc = grow(&b_.cvals_b, &have, ncp))) return rc;
    n_cellpairs_ = ncp;
  }
  b_.cell_cap = (uint32_t)std::min<size_t>(n_cellpairs_, 0xFFFFFFF0u);
  if (L.total_cells > n_cells_ || !b_.cell_offsets) {
    have = n_cells_ * 4;
    if ((rc = grow(&b_.cell_offsets, &have, (size_t)L.total_cells * 4))) return rc;
    n_cells_ = L.total_cells;
  }
  b_.cell_counts = b_.cell_offsets + n_cells_;
  b_.cell_nchunk = b_.cell_offsets + 2 * n_cells_;
  b_.cell_chunk0 = b_.cell_offsets + 3 * n_cells_;
  const size_t max_chunks = (n_cellpairs_ + kCellChunk - 1) / kCellChunk + L.total_cells;
  if ((rc = grow(&b_.chunk_cnt, &n_chunkcnt_, max_chunks * 64))) return rc;
  b_.max_chunks = (uint32_t)(n_chunkcnt_ / 64);
  const size_t scan4 = scan_u4_tmp_words(L.splat_blocks) + 2 * (size_t)L.total_tiles + scan_u4_tmp_words(L.total_tiles);
  if ((rc = grow(&b_.scan_tmp, &n_scan4_, scan4))) return rc;
  // pair capacity
  uint64_t cap = ref_caps_ ? std::max<uint64_t>(ref_cap_, 1) : std::max<uint64_t>({min_cap_, want_cap_, 1});
  cap = std::min<uint64_t>(cap, 0xFFFFFFF0ull);
  if (cap > n_cap_ || !b_.entries) {
    have = n_cap_;
    if ((rc = grow(&b_.entries, &have, cap))) return rc;
    have = n_cap_;
    if ((rc = grow(&b_.scratch, &have, cap))) return rc;
    have = n_cap_;
    if ((rc = grow(&b_.qkey, &have, cap))) return rc;
    have = n_cap_;
    if ((rc = grow(&b_.qkey_s, &have, cap))) return rc;
    n_cap_ = cap;
    n_debug_ = 0;  // debug arrays follow the capacity
  }
  b_.cap = ref_caps_ ? ref_cap_ : (uint32_t)n_cap_;  // under the caps: the drop threshold itself
  if (debug) {
    if (n_debug_ < n_cap_ || !b_.keys20) {
      have = n_debug_;
      if ((rc = grow(&b_.keys20, &have, n_cap_))) return rc;
      have = n_debug_;
      if ((rc = grow(&b_.entries_canon, &have, n_cap_))) return rc;
      n_debug_ = n_cap_;
    }
  }
  const size_t rblocks = (n_cellpairs_ + kRadixBlockItems - 1) / kRadixBlockItems;
  if ((rc = grow(&b_.radix_hist, &n_hist_, rblocks * kRadixBuckets))) return rc;
  const size_t tmpw = std::max({scan_u32_tmp_words((uint32_t)(rblocks * kRadixBuckets)),
                                scan_u32_tmp_words(L.total_cells), scan_u32_tmp_words(L.total_morton)});
  if ((rc = grow(&b_.radix_tmp, &n_hist_tmp_, tmpw))) return rc;
  if (L.total_morton > n_morton_ || !b_.tile_offsets) {
    have = n_morton_;
    if ((rc = grow(&b_.tile_offsets, &have, L.total_morton))) return rc;
    have = n_morton_;
    if ((rc = grow(&b_.tile_counts, &have, L.total_morton))) return rc;
    have = n_morton_;
    if ((rc = grow(&b_.tile_large, &have, L.total_morton))) return rc;
    have = n_morton_ * 3;
    if ((rc = grow(&b_.tile_dmin, &have, (size_t)L.total_morton * 3))) return rc;
    n_morton_ = L.total_morton;
  }
  b_.tile_dmax = b_.tile_dmin + n_morton_;
  b_.tile_unl = ref_caps_ ? b_.tile_dmin + 2 * n_morton_ : nullptr;  // per-tile unlisted large splats
  b_.ref_caps = ref_caps_ ? 1u : 0u;
  if (L.total_tiles > n_tiles_ || !b_.small_list) {
    have = n_tiles_;
    if ((rc = grow(&b_.small_list, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.large_list_t, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.large_bases, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.large_tiles, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.pend_tiles, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.pend_large, &have, L.total_tiles))) return rc;
    have = n_tiles_;
    if ((rc = grow(&b_.large_order, &have, L.total_tiles))) return rc;
    have = n_tiles_ * 2;
    if ((rc = grow(&b_.pend_aux, &have, (size_t)L.total_tiles * 2))) return rc;
    have = n_tiles_ * 256 * 6;
    if ((rc = grow(&b_.px_state, &have, (size_t)L.total_tiles * 256 * 6))) return rc;
    n_tiles_ = L.total_tiles;
  }
  // first chunks of the deferred frames' large tiles: a large tile holds more than 4096 of the at
  // most 'cap' stored entries
  {
    const size_t ft = std::min<size_t>(L.total_tiles, (size_t)b_.cap / (HDSR_MAX_TILE_ENTRIES + 1) + 1);
    if (ft > n_first_ || !b_.first) {
      have = n_first_ * HDSR_MAX_TILE_ENTRIES;
      if ((rc = grow(&b_.first, &have, ft * HDSR_MAX_TILE_ENTRIES))) return rc;
      n_first_ = ft;
    }
    b_.first_tiles = (uint32_t)n_first_;
  }
  // segments: every small tile, plus at most one 4096-entry chunk per 4096 pairs and per large tile
  const size_t segs = 2 * (size_t)L.total_tiles + (n_cap_ + HDSR_MAX_TILE_ENTRIES - 1) / HDSR_MAX_TILE_ENTRIES + 1;
  size_t have_pend = n_seg_;  // pend_segs[] has the capacity of segs[]
  if ((rc = grow(&b_.segs, &n_seg_, segs))) return rc;
  if ((rc = grow(&b_.pend_segs, &have_pend, segs))) return rc;
  b_.seg_cap = (uint32_t)n_seg_;
  if (!b_.stats) {
    size_t h = 0;
    if ((rc = grow(&b_.stats, &h, 1))) return rc;
  }
  // the frame table and its three start arrays (uint32, num_frames + 1 each) share one allocation,
  // laid out as the host stages them, so a batch's upload is one copy
  const size_t slots = L.num_frames + (12 * ((size_t)L.num_frames + 1) + sizeof(FrameDesc) - 1) / sizeof(FrameDesc);
  if (slots > n_frames_ || !d_frames_) {
    have = n_frames_;
    if ((rc = grow(&d_frames_, &have, slots))) return rc;
    n_frames_ = slots;
  }
  d_starts_ = reinterpret_cast<uint32_t*>(d_frames_ + L.num_frames);
  if (!hstats_) {
    if (hipHostMalloc((void**)&hstats_, sizeof(DevStats), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      err = "pinned host allocation failed";
      return HDSR_E_OOM;
    }
    memset(hstats_, 0, sizeof(DevStats));
  }
  if (!done_ev_ && hipEventCreateWithFlags(&done_ev_, hipEventDisableTiming) != hipSuccess) {
    err = "hipEventCreate failed";
    return HDSR_E_HIP;
  }

  const size_t sb = sizeof(FrameDesc) * L.num_frames + 12 * (L.num_frames + 1);
  if (sb > stage_bytes_) {
    for (int i = 0; i < kRing; ++i) {
      if (stage_ev_[i]) (void)hipEventSynchronize(stage_ev_[i]);
      if (stage_[i]) (void)hipHostFree(stage_[i]);
      stage_[i] = nullptr;
      if (hipHostMalloc(&stage_[i], sb, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        err = "pinned host allocation failed";
        return HDSR_E_OOM;
      }
      if (!stage_ev_[i] && hipEventCreateWithFlags(&stage_ev_[i], hipEventDisableTiming) != hipSuccess) {
        err = "hipEventCreate failed";
        return HDSR_E_HIP;
      }
    }
    stage_bytes_ = sb;
  }
  b_.frames = d_frames_;
  b_.block_frame_start = d_starts_;
  b_.tile_frame_start = d_starts_ + (L.num_frames + 1);
  b_.cell_frame_start = d_starts_ + 2 * (L.num_frames + 1);
  b_.num_frames = L.num_frames;
  b_.splat_blocks = L.splat_blocks;
  b_.total_tiles = L.total_tiles;
  b_.total_morton = L.total_morton;
  b_.total_cells = L.total_cells;
  b_.cell_key_bits = L.cell_key_bits;
  if (!debug) {
    // keep the buffers, but the kernels only write them when asked
  }
  return 0;
}

int Engine::upload_frames(const FrameDesc* frames, const Layout& L, hipStream_t stream) {
  const int slot = stage_next_;
  stage_next_ = (stage_next_ + 1) % kRing;
  if (hipEventSynchronize(stage_ev_[slot]) != hipSuccess) {
    err = "hipEventSynchronize failed";
    return HDSR_E_HIP;
  }
  char* s = static_cast<char*>(stage_[slot]);
  const size_t fb = sizeof(FrameDesc) * L.num_frames;
  memcpy(s, frames, fb);
  uint32_t* st = reinterpret_cast<uint32_t*>(s + fb);
  for (uint32_t v = 0; v <= L.num_frames; ++v) {
    st[v] = L.block_start[v];
    st[L.num_frames + 1 + v] = L.tile_start[v];
    st[2 * (L.num_frames + 1) + v] = L.cell_start[v];
  }
  if (hipMemcpyAsync(d_frames_, s, fb + 12 * (L.num_frames + 1), hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipEventRecord(stage_ev_[slot], stream) != hipSuccess) {
    err = "frame table upload failed";
    return HDSR_E_HIP;
  }
  return 0;
}

void Engine::enqueue_stages(const Bufs& b, hipStream_t stream, bool gbuffer, bool debug, hipEvent_t* ev) {
  launch_clear_stats(b, stream);
  if (ev) (void)hipEventRecord(ev[0], stream);
  {
    Bufs bp = b;
    bp.write_cold = (gbuffer || debug) ? 1u : 0u;  // the G-buffer-only words
    launch_project(bp, stream);
  }
  if (ev) (void)hipEventRecord(ev[1], stream);
  launch_scan_blocks(b, stream);
  if (ev) (void)hipEventRecord(ev[2], stream);
  launch_emit_pairs(b, stream);
  if (ev) (void)hipEventRecord(ev[3], stream);
  launch_group_pairs(b, stream);
  if (ev) (void)hipEventRecord(ev[4], stream);
  launch_tile_ranges(b, stream);
  if (ev) (void)hipEventRecord(ev[5], stream);
  launch_classify(b, stream);
  if (ev) (void)hipEventRecord(ev[6], stream);
  if (debug && b.entries_canon)
    (void)hipMemcpyAsync(b.entries_canon, b.entries, (size_t)b.cap * 4, hipMemcpyDeviceToDevice, stream);
  // Deferred chunk sort when nothing reads the whole lists: no debug readback, no G-buffer (its
  // votes are not part of the saved pixel state), no backward pass.
  deferred_ = !debug && !gbuffer && !b.drec && !b.ref_caps;
  last_deferred_ = deferred_;
  const bool first = deferred_ && first_chunk_;
  last_first_ = first;
  Bufs r1 = b, r2 = b;
  r1.round = deferred_ ? 1u : 0u;
  r2.round = 2u;
  r1.first_chunks = r2.first_chunks = first ? 1u : 0u;
  if (first) {
    // the large tiles' first chunks and the small tiles' sorts in one launch (the sort stage of
    // round 1 is in the bucket stage's time)
    launch_first_chunks(r1, stream);
    if (ev) (void)hipEventRecord(ev[7], stream);
  } else {
    launch_bucket(r1, stream);
    if (ev) (void)hipEventRecord(ev[7], stream);
    launch_segment_sort(r1, stream);
  }
  // Stage events: sort = [7, 8] + [10, 11], composite = [8, 10] + [11, 9] (the rounds interleave).
  if (ev) (void)hipEventRecord(ev[8], stream);
  if (deferred_) {
    launch_composite(r1, stream, gbuffer, f16_);
    if (ev) (void)hipEventRecord(ev[kStages + 1], stream);
    if (first) launch_bucket(r2, stream);  // the bucket order of the tiles left unsaturated
    launch_segment_sort(r2, stream);
    if (ev) (void)hipEventRecord(ev[kStages + 2], stream);
    launch_composite(r2, stream, gbuffer, f16_);
  } else {
    if (ev) (void)hipEventRecord(ev[kStages + 1], stream);
    if (ev) (void)hipEventRecord(ev[kStages + 2], stream);
    launch_composite(r1, stream, gbuffer, f16_);
  }
  if (ev) (void)hipEventRecord(ev[9], stream);
}

void Engine::drop_graphs() {
  for (GraphEntry& g : graphs_) (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
}

int Engine::run_graph(const FrameDesc* frames_in, uint32_t V, hipStream_t stream, bool gbuffer, bool* replayed) {
  *replayed = false;
  if (V == 0 || V > kMaxFrames) {
    err = "frame count out of range";
    return HDSR_E_INVALID_ARG;
  }
  std::vector<FrameDesc> frames(frames_in, frames_in + V);
  Layout L = layout(frames.data(), V);
  const uint64_t epoch0 = epoch_;
  int rc = ensure(L, false);
  if (rc) return rc;
  lay_ = L;
  const size_t fb = sizeof(FrameDesc) * V, sb = 12 * (size_t)(V + 1);
  if (fb + sb > gstage_bytes_) {
    if (gstage_) (void)hipHostFree(gstage_);
    gstage_ = nullptr;
    gstage_bytes_ = 0;
    if (hipHostMalloc(&gstage_, fb + sb, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      err = "pinned host allocation failed";
      return HDSR_E_OOM;
    }
    gstage_bytes_ = fb + sb;
    ++epoch_;
  }
  if (epoch_ != epoch0) drop_graphs();
  // the previous replay on this engine must be done reading the staging buffer
  if (done_ev_ && hipEventSynchronize(done_ev_) != hipSuccess) {
    err = "hipEventSynchronize failed";
    return HDSR_E_HIP;
  }
  char* st = static_cast<char*>(gstage_);
  memcpy(st, frames.data(), fb);
  uint32_t* tab = reinterpret_cast<uint32_t*>(st + fb);
  for (uint32_t v = 0; v <= V; ++v) {
    tab[v] = L.block_start[v];
    tab[V + 1 + v] = L.tile_start[v];
    tab[2 * (V + 1) + v] = L.cell_start[v];
  }
  const uint64_t sig[7] = {V, L.splat_blocks, L.total_tiles, L.total_morton, L.total_cells, L.cell_key_bits,
                           (uint64_t)gbuffer | ((uint64_t)L.packed << 1) | ((uint64_t)bucket_wide_ << 2) |
                               ((uint64_t)first_chunk_ << 3)};
  hipGraphExec_t exec = nullptr;
  for (const GraphEntry& g : graphs_)
    if (memcmp(g.sig, sig, sizeof sig) == 0) exec = g.exec;
  if (!exec) {
    // capture costs milliseconds: a shape is captured only once it has been seen before
    bool seen = false;
    for (const auto& s : seen_sigs_) seen |= memcmp(s.data(), sig, sizeof sig) == 0;
    if (!seen) {
      std::array<uint64_t, 7> a;
      memcpy(a.data(), sig, sizeof sig);
      if (seen_sigs_.size() >= 1024) seen_sigs_.erase(seen_sigs_.begin());
      seen_sigs_.push_back(a);
      return run(frames_in, V, stream, gbuffer, false);
    }
  }
  if (!exec) {
    Bufs b = b_;
    b.bucket_wide = bucket_wide_;
    b.keys20 = nullptr;
    b.entries_canon = nullptr;
    b.packed = L.packed;
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      err = "hipStreamBeginCapture failed";
      return HDSR_E_HIP;
    }
    (void)hipMemcpyAsync(d_frames_, st, fb + sb, hipMemcpyHostToDevice, stream);
    enqueue_stages(b, stream, gbuffer, false, nullptr);
    (void)hipMemcpyAsync(hstats_, b.stats, sizeof(DevStats), hipMemcpyDeviceToHost, stream);
    if (hipStreamEndCapture(stream, &graph) != hipSuccess || !graph) {
      (void)hipGetLastError();
      err = "graph capture failed";
      return HDSR_E_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      (void)hipGetLastError();
      err = "hipGraphInstantiate failed";
      return HDSR_E_HIP;
    }
    if (graphs_.size() >= 128) {
      (void)hipGraphExecDestroy(graphs_.front().exec);
      graphs_.erase(graphs_.begin());
    }
    GraphEntry e;
    memcpy(e.sig, sig, sizeof sig);
    e.exec = exec;
    graphs_.push_back(e);
  } else {
    *replayed = true;
  }
  if (hipGraphLaunch(exec, stream) != hipSuccess || hipEventRecord(done_ev_, stream) != hipSuccess) {
    err = std::string("graph launch failed: ") + hipGetErrorString(hipGetLastError());
    return HDSR_E_HIP;
  }
  pending_ = true;
  return 0;
}

int Engine::backward(uint32_t n, const float* grad_rgba, const DiffGrads& g, hipStream_t stream) {
  if (!diff_ || !b_.drec || !b_.gacc || lay_.num_frames != 1) {
    err = "backward needs a differentiable-mode render of one frame";
    return HDSR_E_STATE;
  }
  launch_backward(b_, n, grad_rgba, g, stream);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    err = std::string("backward launch failed: ") + hipGetErrorString(e);
    return HDSR_E_HIP;
  }
  return 0;
}

int Engine::run(const FrameDesc* frames_in, uint32_t V, hipStream_t stream, bool gbuffer, bool debug) {
  if (V == 0 || V > kMaxFrames) {
    err = "frame count out of range";
    return HDSR_E_INVALID_ARG;
  }
  std::vector<FrameDesc> frames(frames_in, frames_in + V);
  Layout L = layout(frames.data(), V);
  int rc = ensure(L, debug);
  if (rc) return rc;
  lay_ = L;
  if ((rc = upload_frames(frames.data(), L, stream))) return rc;
  Bufs b = b_;
  b.packed = L.packed;
  b.bucket_wide = bucket_wide_;
  if (!debug) {
    b.keys20 = nullptr;
    b.entries_canon = nullptr;
  }
  enqueue_stages(b, stream, gbuffer, debug, profile_ ? prof_slot() : nullptr);
  if (hipMemcpyAsync(hstats_, b.stats, sizeof(DevStats), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipEventRecord(done_ev_, stream) != hipSuccess) {
    err = "stats copy failed";
    return HDSR_E_HIP;
  }
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    err = std::string("kernel launch failed: ") + hipGetErrorString(le);
    return HDSR_E_HIP;
  }
  pending_ = true;
  return 0;
}

int Engine::composite_only(const FrameDesc* frames_in, uint32_t V, hipStream_t stream, bool gbuffer) {
  if (!pending_ && lay_.num_frames == 0) {
    err = "composite_only before any render";
    return HDSR_E_STATE;
  }
  if (V != lay_.num_frames) {
    err = "composite_only needs the frames of the last render";
    return HDSR_E_STATE;
  }
  std::vector<FrameDesc> frames(frames_in, frames_in + V);
  Layout L = layout(frames.data(), V);
  int rc = upload_frames(frames.data(), L, stream);
  if (rc) return rc;
  if (deferred_ && gbuffer) {  // the G-buffer reads whole lists; a deferred frame sorted only a part
    err = "composite_only with the G-buffer after a frame rendered without it";
    return HDSR_E_STATE;
  }
  if (deferred_) {
    // the lists as the last render left them: the rounds meet the same saturation again, and
    // round 1 lists the same tiles for round 2 (whose chunks that render sorted)
    Bufs r1 = b_, r2 = b_;
    r1.round = 1;
    r2.round = 2;
    r1.first_chunks = r2.first_chunks = last_first_ ? 1u : 0u;
    launch_reset_pending(b_, stream);
    launch_composite(r1, stream, gbuffer, f16_);
    launch_composite(r2, stream, gbuffer, f16_);
  } else {
    launch_reset_pending(b_, stream);  // the rasterizer's counters count this run
    launch_composite(b_, stream, gbuffer, f16_);
  }
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    err = std::string("kernel launch failed: ") + hipGetErrorString(le);
    return HDSR_E_HIP;
  }
  return 0;
}

int Engine::complete_lists(hipStream_t stream) {
  if (!deferred_) return 0;
  Bufs r3 = b_;
  r3.round = 3;
  r3.first_chunks = last_first_ ? 1u : 0u;
  if (last_first_) launch_bucket(r3, stream);  // the tiles round 2 did not order
  launch_segment_sort(r3, stream);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    err = std::string("kernel launch failed: ") + hipGetErrorString(le);
    return HDSR_E_HIP;
  }
  deferred_ = false;
  return 0;
}

hipEvent_t* Engine::prof_slot() {
  if (prof_ev_.empty()) {
    prof_ev_.resize((size_t)kProfRing * kProfEvents);
    for (hipEvent_t& e : prof_ev_)
      if (hipEventCreate(&e) != hipSuccess) {
        prof_ev_.clear();
        profile_ = false;
        return nullptr;
      }
  }
  if (prof_count_ == kProfRing) prof_harvest_one();
  const int slot = (prof_head_ + prof_count_) % kProfRing;
  ++prof_count_;
  return &prof_ev_[(size_t)slot * kProfEvents];
}

void Engine::prof_harvest_one() {
  hipEvent_t* ev = &prof_ev_[(size_t)prof_head_ * kProfEvents];
  (void)hipEventSynchronize(ev[kStages]);
  auto span = [&](int a, int b) {
    float ms = 0.0f;
    return hipEventElapsedTime(&ms, ev[a], ev[b]) == hipSuccess ? (double)ms : 0.0;
  };
  for (int s = 0; s < kStages - 2; ++s) prof_acc_[s] += span(s, s + 1);
  const int sort = kStages - 2, comp = kStages - 1, e1 = kStages + 1, e2 = kStages + 2;
  prof_acc_[sort] += span(sort, sort + 1) + span(e1, e2);
  prof_acc_[comp] += span(comp, e1) + span(e2, kStages);
  ++prof_runs_;
  prof_head_ = (prof_head_ + 1) % kProfRing;
  --prof_count_;
}

int Engine::stage_times(float* ms, uint32_t* runs) {
  while (prof_count_ > 0) prof_harvest_one();
  for (int s = 0; s < kStages; ++s) {
    ms[s] = prof_runs_ ? (float)(prof_acc_[s] / prof_runs_) : 0.0f;
    prof_acc_[s] = 0.0;
  }
  *runs = prof_runs_;
  prof_runs_ = 0;
  return 0;
}

int Engine::wait() {
  if (!done_ev_) return 0;
  if (hipEventSynchronize(done_ev_) != hipSuccess) {
    err = std::string("device error: ") + hipGetErrorString(hipGetLastError());
    return HDSR_E_HIP;
  }
  // the bucket cascade's workgroup size for the next frames, from this frame's large tiles: both
  // sizes give the same order; 512 threads are faster for the usual tiles of a few thousand
  // entries, 1024 when tiles of 100k entries would leave one 512-thread workgroup as the tail
  const DevStats& s = *hstats_;
  bucket_wide_ = s.large_tiles != 0 && s.scratch_used / s.large_tiles > 16384u;
  // first chunks pay when round 2 has little to scatter: the resumed tiles hold less than half of
  // the large tiles' references and none is so long that its scatter, one workgroup after round 1,
  // would form the frame's tail
  if (last_deferred_ && !bucket_scatter_)
    first_chunk_ = (uint64_t)s.resumed_refs * 2u < s.scratch_used && s.resumed_max <= 32768u;
  return 0;
}

int Engine::grow_for_overflow() {
  const uint64_t need = (uint64_t)hstats_->small_pairs + hstats_->large_pairs;
  want_cap_ = need + need / 4 + 4096;
  return 0;
}

}  // namespace hdsr
