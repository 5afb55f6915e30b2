// engine.h - host-side owner of one batch's device buffers and the stage sequence.
//
// An Engine renders a batch of frames (FrameDesc, one per view) with one launch per stage.  It
// keeps its buffers between calls and only grows them.  The pair capacity is checked on the
// device: a batch whose tile references exceed it is flagged in DevStats and must be run again
// after grow_for_overflow() (the plain context does that in hdsr_sync).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "hdsr.h"
#include "hdsr_internal.h"

namespace hdsr {

struct Layout {
  uint32_t num_frames = 0;
  uint32_t splat_blocks = 0;
  uint32_t total_tiles = 0;
  uint32_t total_morton = 0;
  uint32_t total_cells = 0;
  uint32_t cell_key_bits = 0;
  uint32_t packed = 0;                // every frame has packed splat records
  std::vector<uint32_t> block_start;  // [V+1]
  std::vector<uint32_t> tile_start;   // [V+1]
  std::vector<uint32_t> cell_start;   // [V+1]
};

class Engine {
 public:
  Engine() = default;
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Fills the batch-layout fields of frames[0..V) and computes the layout.
  static Layout layout(FrameDesc* frames, uint32_t V);

  // Enqueues the full path for 'frames' on 'stream'.  Returns 0 or an HDSR_E_* code.
  int run(const FrameDesc* frames, uint32_t V, hipStream_t stream, bool gbuffer, bool debug);
  // Re-runs only the rasterizer on the previous run's tile lists (same frames, new outputs ok).
  int composite_only(const FrameDesc* frames, uint32_t V, hipStream_t stream, bool gbuffer);
  // As run(), through a hipGraph: the stage sequence is captured once per batch shape (the
  // totals that set grid sizes) and replayed afterwards.  Per-frame data (scene and output
  // pointers, cameras, block / tile tables) lives in the device frame table, which every replay
  // re-uploads from a fixed pinned buffer, so one graph serves any mix with the same totals.
  // *replayed tells whether a cached graph was used.  The stream must not be the null stream.
  int run_graph(const FrameDesc* frames, uint32_t V, hipStream_t stream, bool gbuffer, bool* replayed);

  // Host copy of the device counters of the last run (valid after wait()).
  const DevStats& host_stats() const { return *hstats_; }
  // whether the last enqueued frame ran the deferred chunk sort (first chunks, then the resumed
  // tiles' other chunks) rather than sorting every chunk
  bool last_frame_deferred() const { return last_deferred_; }
  bool last_frame_first_chunks() const { return last_first_; }
  int wait();
  bool overflowed() const { return hstats_ && hstats_->overflow != 0; }
  // Grows the pair capacity to hold the last run's references (with headroom).
  int grow_for_overflow();
  void set_min_capacity(uint64_t cap) { min_cap_ = cap; }
  // The reference's WebGPU caps (HDSR_REF_CAPS): the pair capacity is exactly maxTotalTileRefs,
  // references past it are dropped rather than grown for, at most 65,535 large splats are stored
  // and chunks past the first 1,024 keep their bucket order.  One frame per run.
  void set_ref_caps(uint32_t max_total_tile_refs) {
    ref_caps_ = true;
    ref_cap_ = max_total_tile_refs;
  }
  // Differentiable mode: keeps the fp32 record table and the per-splat gradient slots.
  void set_diff(bool on) { diff_ = on; }
  // Composite with binary16 transmittance / colour (the reference's f16 shader variant).
  void set_f16(bool on) { f16_ = on; }
  // Backward of the last run (one frame, n splats); see launch_backward.
  int backward(uint32_t n, const float* grad_rgba, const DiffGrads& g, hipStream_t stream);

  // Per-stage HIP events (one set per run, kept in a ring until harvested).
  void set_profile(bool on) { profile_ = on; }
  void set_bucket_scatter(bool on) {
    bucket_scatter_ = on;
    first_chunk_ = !on;
  }
  int stage_times(float* ms, uint32_t* runs);
  // Sorts the chunks a deferred frame left unsorted (no-op otherwise), so that the entry lists
  // are whole: for readers of the lists (debug readback).
  int complete_lists(hipStream_t stream);

  const Bufs& bufs() const { return b_; }
  const Layout& last_layout() const { return lay_; }
  bool debug_buffers() const { return b_.keys20 != nullptr; }
  std::string err;

 private:
  int ensure(const Layout& L, bool debug);
  int upload_frames(const FrameDesc* frames, const Layout& L, hipStream_t stream);
  void enqueue_stages(const Bufs& b, hipStream_t stream, bool gbuffer, bool debug, hipEvent_t* ev);
  void drop_graphs();
  struct GraphEntry {
    uint64_t sig[7];
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs_;
  std::vector<std::array<uint64_t, 7>> seen_sigs_;
  uint64_t epoch_ = 0;  // bumped by every (re)allocation: graphs bake buffer addresses
  void* gstage_ = nullptr;
  size_t gstage_bytes_ = 0;
  void release();
  template <typename T>
  int grow(T** p, size_t* have, size_t need);

  Bufs b_{};
  Layout lay_;
  uint64_t min_cap_ = 1u << 20;
  uint64_t want_cap_ = 0;
  bool ref_caps_ = false;
  uint32_t ref_cap_ = 0;
  bool diff_ = false;
  bool f16_ = false;
  bool deferred_ = false;     // the last enqueued frame left chunks unsorted (deferred chunk sort)
  bool last_deferred_ = false;  // as deferred_, but not cleared when the lists are completed
  bool bucket_wide_ = false;  // the last settled frame's large tiles averaged more than 16k entries
  // the last settled frame resumed less than half of its large-tile references in round 2: the next
  // deferred frames select first chunks without the bucket scatter
  bool first_chunk_ = true;
  bool bucket_scatter_ = false;  // HDSR_BUCKET_SCATTER: never select first chunks
  bool last_first_ = false;  // the mode of the last deferred frame (its rounds 2 / 3 follow it)
  size_t n_diff_ = 0;
  // allocation sizes (elements)
  size_t n_splat_ = 0, n_blocks_ = 0, n_scan4_ = 0, n_cap_ = 0, n_hist_ = 0, n_hist_tmp_ = 0;
  size_t n_morton_ = 0, n_tiles_ = 0, n_seg_ = 0, n_frames_ = 0, n_debug_ = 0;
  size_t n_cellpairs_ = 0, n_cells_ = 0, n_chunkcnt_ = 0, n_radix_tmp_need_ = 0;
  size_t n_first_ = 0;  // large tiles first[] holds
  FrameDesc* d_frames_ = nullptr;
  uint32_t* d_starts_ = nullptr;
  // pinned staging ring for the frame table
  static constexpr int kRing = 4;
  void* stage_[kRing] = {};
  size_t stage_bytes_ = 0;
  hipEvent_t stage_ev_[kRing] = {};
  int stage_next_ = 0;
  DevStats* hstats_ = nullptr;
  hipEvent_t done_ev_ = nullptr;
  bool pending_ = false;
  // profiling
  static constexpr int kStages = HDSR_STAGE_COUNT;
  static constexpr int kProfRing = 128;
  static constexpr int kProfEvents = kStages + 3;  // stage boundaries + the two of the deferred rounds
  bool profile_ = false;
  std::vector<hipEvent_t> prof_ev_;  // kProfRing * (kStages + 1)
  int prof_head_ = 0, prof_count_ = 0;
  double prof_acc_[kStages] = {};
  uint32_t prof_runs_ = 0;
  hipEvent_t* prof_slot();
  void prof_harvest_one();
};

// Parses the 256-byte uniform block into a FrameDesc's camera fields
// (layout of gpu_project_splats.wgsl:27-44).
void parse_uniforms(const void* p, FrameDesc* f);

}  // namespace hdsr
