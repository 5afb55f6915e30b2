/* hdsr.h - C ABI of the MI355X splat compositor (drop-in for the reference's gpuTileSplatCompositor).
 *
 * Plain pointers and sizes only; no HIP or torch types cross this boundary (a stream is passed
 * as void*, NULL = the default stream).  Every function returns 0 on success or a negative
 * HDSR_E_* code; hdsr_last_error() gives the message of the last failure on a context.  A context
 * is bound to the device that was current when it was created and must be used from one host
 * thread at a time.
 *
 * Each entry point names the reference interface it replaces (src/gpuTileSplatCompositor.ts
 * unless stated).  The JavaScript facade (node/gpuTileSplatCompositor.js) and the Python facade
 * (hdsr/compositor.py) bind exactly these symbols.
 */
#ifndef HDSR_H
#define HDSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDSR_OK 0
#define HDSR_E_INVALID_ARG (-1)
#define HDSR_E_OOM (-2)
#define HDSR_E_CAPACITY (-3)
#define HDSR_E_HIP (-4)
#define HDSR_E_STATE (-5)
#define HDSR_E_UNSUPPORTED (-6)

/* TILE_SPLAT_FRAME_UNIFORM_BYTES (gpuTileSplatCompositor.ts:18) */
#define HDSR_FRAME_UNIFORM_BYTES 256

/* planTileSplatCompositor input (gpuTileSplatCompositor.ts:109-116). */
typedef struct hdsr_plan_in {
  uint32_t viewport_width;
  uint32_t viewport_height;
  uint32_t tile_size_px;          /* must be 16 (the rasterizer's tile) */
  uint32_t splat_count;
  float tile_entry_multiplier;    /* <= 0 means the reference default 16 */
} hdsr_plan_in;

/* TileSplatCompositorPlan (gpuTileSplatCompositor.ts:39-49). */
typedef struct hdsr_plan {
  uint32_t viewport_width;
  uint32_t viewport_height;
  uint32_t tile_size_px;
  uint32_t tile_columns;
  uint32_t tile_rows;
  uint32_t tile_count;
  uint32_t morton_tile_count;
  uint32_t splat_count;
  uint32_t max_total_tile_refs;   /* reference budget, reported; capacities here are lifted */
} hdsr_plan;

/* planTileSplatCompositor (gpuTileSplatCompositor.ts:109-139): identical arithmetic. */
int hdsr_plan_compute(const hdsr_plan_in* in, hdsr_plan* out);

/* createTileSplatCompositor options (gpuTileSplatCompositor.ts:141-145) and output selection. */
#define HDSR_F16_COMPOSITE 1u   /* options.f16: composite with binary16 transmittance / colour
                                   (gpu_tile_splat_composite_f16.wgsl); not combinable with
                                   HDSR_DIFFERENTIABLE */
#define HDSR_REF_CAPS 2u        /* the reference's WebGPU caps instead of lifted capacities:
                                   references past plan.max_total_tile_refs are dropped
                                   (gpu_tile_splat_composite.wgsl:114), only the first 65,535
                                   large splats are scattered (gpu_classify_large_splats.wgsl:52,62;
                                   the rest keep their slots unwritten) and chunks past the first
                                   1,024 keep their bucket order (gpu_tile_bucket_sort.wgsl:126,138).
                                   Each tile keeps the entries the reference stores; where the
                                   reference would read slots it never wrote, nothing is read.
                                   No capacity re-run; not combinable with HDSR_DIFFERENTIABLE */
#define HDSR_GBUFFER 4u         /* also write the oct-normal and material/emissive/confidence words */
#define HDSR_DEBUG 8u           /* keep the pre-sort entry list and the 20-bit keys for readback */
#define HDSR_OUT_RGBA16 16u     /* also write binary16 RGBA (the reference's rgba16float texture) */
#define HDSR_PROFILE 32u        /* record a HIP event between pipeline stages of every render */
#define HDSR_DIFFERENTIABLE 64u /* differentiable mode (see hdsr_backward): opacities are logits,
                                   the rasterizer reads fp32 colour / opacity (no f16 cache
                                   rounding) and keeps per-pixel state for the backward; not
                                   combinable with HDSR_GBUFFER or HDSR_F16_COMPOSITE */
#define HDSR_BUCKET_SCATTER 128u /* frames that sort only the chunks their pixels reach keep the
                                   whole bucket scatter of the large tiles in their first round
                                   (by default a frame after one that resumed less than half of its
                                   large-tile references selects each large tile's first chunk
                                   without it); lists, order and pixels are the same either way */

typedef struct hdsr_ctx hdsr_ctx;

/* createTileSplatCompositor (gpuTileSplatCompositor.ts:141-728) / resources.destroy() (:700-726). */
int hdsr_create(const hdsr_plan* plan, uint32_t flags, hdsr_ctx** out);
void hdsr_destroy(hdsr_ctx* ctx);

#define HDSR_MEM_HOST 0u        /* pointers are host memory: the context keeps device copies */
#define HDSR_MEM_DEVICE 1u      /* pointers are device memory: borrowed, must outlive the context */

/* The splat storage buffers bound by createTileSplatBindGroups (gpuTileSplatCompositor.ts:792-995),
 * in the layouts of uploadSplatAttributeBuffers (src/splats.ts:376-500):
 *   positions 3N f32, scales 3N f32 (log-scale), rotations 4N f32 (w,x,y,z), opacities N f32,
 *   materials N u32 (pack2x16float(roughness, metalness)), normals: normals_len f32
 *   ([geometric 3N | detail 3N], or 0 = derive from the covariance), sh_data N*(3+3K+3) f32
 *   (dc rgb, K = (d+1)^2-1 coefficients rgb-interleaved, emissive rgb),
 *   sorted_indices N u32 (a permutation; the identity is allowed).
 * The reference binds shData without a degree and reads the degree from the frame uniforms
 * (u32 word 25, gpu_project_splats.wgsl:294-296; the bind call src/splatRenderer.ts:946 names no
 * degree).  Here sh_degree = HDSR_SH_DEGREE_AUTO takes it from sh_data_len: N*(6+3K) floats
 * must match exactly one degree 0..3 (6, 15, 30 or 51 floats per splat).  With an explicit
 * sh_degree, a nonzero sh_data_len must hold at least N*(6+3K) floats.  Either way the degree of
 * every frame's uniforms (hdsr_set_frame) must equal the bound degree, or the call fails with
 * HDSR_E_INVALID_ARG: a WebGPU host with mismatched buffers would read other splats' rows. */
#define HDSR_SH_DEGREE_AUTO 0xFFFFFFFFu
typedef struct hdsr_splats {
  const float* positions;
  const float* scales;
  const float* rotations;
  const float* opacities;
  const uint32_t* materials;
  const float* normals;
  uint64_t normals_len;
  const float* sh_data;
  const uint32_t* sorted_indices;
  uint32_t count;
  uint32_t sh_degree;             /* 0..3 or HDSR_SH_DEGREE_AUTO */
  uint32_t memory;                /* HDSR_MEM_HOST or HDSR_MEM_DEVICE */
  uint64_t sh_data_len;           /* floats in sh_data; 0 = not stated (needs an explicit degree) */
} hdsr_splats;

/* createTileSplatBindGroups (gpuTileSplatCompositor.ts:792-995). */
int hdsr_bind_splats(hdsr_ctx* ctx, const hdsr_splats* splats);

/* Replaces the sortedIndices buffer contents (the reference's encodeSort output,
 * src/gpuSortPrototype.ts:275-287 / src/splatSort.ts:32-55). */
int hdsr_update_sorted_indices(hdsr_ctx* ctx, const uint32_t* sorted_indices, uint32_t memory);

/* queue.writeBuffer(frameUniformBuffer, ...) of writeTileSplatFrameUniforms
 * (gpuTileSplatCompositor.ts:730-774; src/splatRenderer.ts:1030): the 256-byte block verbatim.
 * Its tile grid and splatCount must equal the plan's, and its shDegree (word 25) the bound
 * splats' degree (checked here once splats are bound, else by hdsr_bind_splats and hdsr_render). */
int hdsr_set_frame(hdsr_ctx* ctx, const void* uniforms256);

/* The sortedIndices producer on the GPU: createViewDepthSortInput + encodeGpuSortPrototype
 * (src/gpuSortPrototype.ts:120-140, 275-287).  Sorts the bound splats back to front for the
 * current frame's view matrix (key f32(-view z), view z evaluated in double; descending, ties to the
 * lower id) and binds the result as the context's own sortedIndices.  Enqueued on 'stream'; a
 * following hdsr_render on the same stream uses it.  Needs hdsr_bind_splats and hdsr_set_frame. */
int hdsr_sort_back_to_front(hdsr_ctx* ctx, void* stream);

/* Allocation-free variant of hdsr_sort_back_to_front.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * hosts that forbid allocation).  Same network and result as hdsr_sort_back_to_front, for
 * 'count' splats whose positions (3*count f32) are in device memory, with the view matrix of the
 * 256-byte frame uniforms; writes count sort ranks' splat ids to the device buffer sorted_out.
 * Needs no context and calls no allocator: its scratch is a fixed device array of
 * HDSR_NOALLOC_MAX_SPLATS pairs per GPU, held until the process ends.  count above that returns
 * HDSR_E_CAPACITY.  Calls that overlap on one GPU (other streams or threads) share that array
 * and race.  Enqueued on 'stream'; errors are reported by hdsr_last_error(NULL). */
#define HDSR_NOALLOC_MAX_SPLATS 4194304u
int hdsr_sort_back_to_front_noalloc(const float* positions, uint32_t count, const void* uniforms256,
                                    uint32_t* sorted_out, void* stream);

/* encodeFullComputeCompositorPipeline (gpuTileSplatCompositor.ts:1023-1179): enqueues the whole
 * path on 'stream' and returns.  Outputs are complete once the stream reaches this point, unless
 * the pair capacity overflowed; hdsr_sync() detects that, grows the capacity and re-renders. */
int hdsr_render(hdsr_ctx* ctx, void* stream);

/* encodeCompositeOnly (gpuTileSplatCompositor.ts:1001-1014): re-runs only the rasterizer on the
 * last frame's sorted tile lists (static camera). */
int hdsr_composite_only(hdsr_ctx* ctx, void* stream);

/* Waits for the last render, and re-renders it with a larger pair capacity if it overflowed. */
int hdsr_sync(hdsr_ctx* ctx);

/* Output textures (src/splatRenderer.ts:908-943) as device buffers owned by the context. */
typedef struct hdsr_outputs {
  float* rgba;                    /* H*W*4 f32, linear RGB + alpha (row-major, top row first) */
  float* depth;                   /* H*W f32: alpha-weighted NDC depth vote */
  uint32_t* gbuffer_normal;       /* H*W u32 pack2x16float(oct normal), if HDSR_GBUFFER */
  uint32_t* gbuffer_material;     /* H*W*4 u32 (material, emissive rg, emissive b, confidence) */
  uint16_t* rgba16;               /* H*W*4 binary16, if HDSR_OUT_RGBA16 */
  uint32_t width;
  uint32_t height;
  uint16_t* lit;                  /* H*W*4 binary16 of hdsr_deferred_lighting, after its first call */
} hdsr_outputs;
int hdsr_get_outputs(hdsr_ctx* ctx, hdsr_outputs* out);

/* The outputTexture / gbufferTextures.depth arguments of createTileSplatBindGroups
 * (gpuTileSplatCompositor.ts:805-806): later renders write RGBA (H*W*4 f32) and depth (H*W f32)
 * into these caller-owned device buffers instead of the context's own.  NULL for either restores
 * the context's buffer.  The buffers must outlive their use by renders. */
int hdsr_bind_outputs(hdsr_ctx* ctx, float* rgba, float* depth);

#define HDSR_OUT_RGBA 0
#define HDSR_OUT_DEPTH 1
#define HDSR_OUT_GB_NORMAL 2
#define HDSR_OUT_GB_MATERIAL 3
#define HDSR_OUT_RGBA16_BUF 4
#define HDSR_OUT_T_FINAL 5        /* H*W f32 final transmittance (HDSR_DIFFERENTIABLE) */
#define HDSR_OUT_LAST_INDEX 6     /* H*W u32 one past the last blended entry of the pixel's tile list */
#define HDSR_OUT_FILTERED_NORMAL 7 /* H*W u32 oct-normal words of hdsr_screen_space_normals (HDSR_GBUFFER) */
#define HDSR_OUT_LIT 8            /* H*W*4 binary16: hdsr_deferred_lighting's output (rgba16float) */
/* Copies an output to host memory (after hdsr_sync). */
int hdsr_readback(hdsr_ctx* ctx, int which, void* host, uint64_t bytes);

/* The first deferred consumers of the G-buffer (src/splatRenderer.ts:1042-1069), over the last
 * render's fp32 depth and oct-normal words (HDSR_GBUFFER).  The inverse view-projection is the
 * reference's mat4Inverse (src/splatRenderer.ts:148-168: double arithmetic, stored to fp32) of the
 * frame's viewProj; the camera position is the frame's.  Asynchronous on 'stream'. */
typedef struct hdsr_normal_params {
  uint32_t recovery_mode;  /* 0 screen-space normals, 1 grazing recovery, 2 source pass-through;
                              the reference passes 0 for scenes without per-splat normals, else 2
                              (src/splatRenderer.ts:1047) */
  float grazing_start;     /* 0.45 (src/splatRenderer.ts:359-360) */
  float grazing_end;       /* 0.80 */
  float spatial_sigma;     /* 3.0  bilateral filter (src/splatRenderer.ts:419-423) */
  float depth_sigma;       /* 0.05 */
  float normal_sigma;      /* 1.5 */
  uint32_t kernel_radius;  /* 3, at most 8 */
} hdsr_normal_params;
void hdsr_normal_params_default(hdsr_normal_params* out);

/* Screen-space normal reconstruction (src/shaders/gpu_screen_space_normals.wgsl): G-buffer depth
 * and normal words -> the context's filtered-normal buffer (HDSR_OUT_FILTERED_NORMAL).  'params'
 * may be null (the reference's values). */
int hdsr_screen_space_normals(hdsr_ctx* ctx, const hdsr_normal_params* params, void* stream);

/* Depth-guided bilateral normal filter (src/shaders/bilateral_normal_filter.wgsl): the
 * filtered-normal buffer and depth -> the G-buffer normal words, which it overwrites, as the
 * reference writes its result back into the G-buffer normal texture. */
int hdsr_bilateral_normal_filter(hdsr_ctx* ctx, const hdsr_normal_params* params, void* stream);

/* Deferred lighting (createDeferredLightingPass().encode, src/splatRenderer.ts:446-548, encoded at
 * :1115-1139; shader src/shaders/gpu_deferred_lighting.wgsl): GGX specular + Lambertian diffuse of
 * one directional light over the G-buffer, ambient * albedo * AO, Reinhard tone mapping, sRGB
 * encoding and additive emissive, into the context's lit buffer (HDSR_OUT_LIT, rgba16float bits).
 * Inputs: the last render's colour (read as its rgba16float output texture holds it), depth,
 * G-buffer normal words (after hdsr_bilateral_normal_filter when the caller runs the normal
 * passes first, as the reference does) and material words.  The path without an environment map
 * is the one built (the reference's 1 x 1 placeholder: env_intensity / env_rotation have no
 * effect).  Numbers are the reference's JavaScript numbers (double), rounded to fp32 where its
 * Float32Array stores them. */
typedef struct hdsr_material_curve {  /* MaterialCurveParams (src/materialCurves.ts:13-17) */
  double contrast;                    /* 1 = identity */
  double brightness;                  /* 0 */
  double gamma;                       /* 1 */
} hdsr_material_curve;
typedef struct hdsr_lighting_params {
  double light_direction[3];          /* direction the light travels (the shader lights along -dir) */
  double light_intensity;
  double ambient_intensity;           /* ambient colour = (a, a, 1.1 a) */
  uint32_t specular_only;
  double emissive_intensity;          /* 3.0 */
  double emissive_threshold;          /* 0.05 */
  double env_intensity;               /* 1.0 */
  double env_rotation;                /* 0.0 */
  double exposure;                    /* 1.0 */
  hdsr_material_curve roughness_curve, metalness_curve, albedo_curve;  /* DEFAULT_CURVE {1, 0, 1} */
  const float* lighting_view_proj;    /* host, 16 floats (params.lightingViewProj), or null: the frame's */
  const float* lighting_camera_position; /* host, 3 floats, or null: the frame's camera position */
  const float* ao;                    /* device, H*W*4 f32: the AO texture's texels (r = visibility), or
                                         null: visibility 1 everywhere */
  const float* bloom;                 /* device, ceil(H/2)*ceil(W/2)*4 f32: the half-resolution bloom
                                         texture's texels, or null: no bloom */
} hdsr_lighting_params;
void hdsr_lighting_params_default(hdsr_lighting_params* out);
int hdsr_deferred_lighting(hdsr_ctx* ctx, const hdsr_lighting_params* params, void* stream);
/* The 176-byte Params block the pass uses (src/splatRenderer.ts:505-523), written from 'params',
 * the inverse view-projection, the camera position and the viewport.  Host only. */
void hdsr_write_lighting_params(const hdsr_lighting_params* params, const float* view_proj_inverse16,
                                const float* camera_position3, uint32_t width, uint32_t height, float out44[44]);
/* A material curve's 256-entry lookup table as the pass reads it: generateLUT
 * (src/materialCurves.ts:31-42) in double, stored to fp32, then to binary16 by uploadLUT's
 * truncating conversion (:45-57), read back as fp32.  Host only. */
void hdsr_material_curve_lut(const hdsr_material_curve* curve, float out256[256]);

/* Intermediates of the last frame, in the reference's layouts (ranks, Morton ids). */
#define HDSR_DBG_PROJCACHE 0      /* N*14 u32 (gpu_project_splats.wgsl:2-16); HDSR_DEBUG or HDSR_GBUFFER */
#define HDSR_DBG_DEPTH 1          /* N u32 depth bits */
#define HDSR_DBG_TILE_COUNTS 2    /* mortonTileCount u32 */
#define HDSR_DBG_TILE_OFFSETS 3   /* mortonTileCount u32 */
#define HDSR_DBG_ENTRIES 4        /* R u32 sort ranks, final per-tile order */
#define HDSR_DBG_ENTRIES_CANON 5  /* R u32 sort ranks before the per-tile sort (HDSR_DEBUG) */
#define HDSR_DBG_SMALL_LIST 6     /* small tiles (<= 4096 entries), Morton ids, row-major order */
#define HDSR_DBG_LARGE_LIST 7     /* large tiles, Morton ids */
#define HDSR_DBG_LARGE_BASES 8    /* overflow base per large tile */
#define HDSR_DBG_CHUNKS 9         /* (start, count) u32 pairs */
#define HDSR_DBG_KEYS20 10        /* R u32 (q20 << 12) | local index (HDSR_DEBUG) */
#define HDSR_DBG_LARGE_SPLATS 11  /* sort ranks of splats covering > 16 tiles */
#define HDSR_DBG_SORTED 12        /* N u32 the bound sortedIndices (e.g. after hdsr_sort_back_to_front) */
int hdsr_debug_size(hdsr_ctx* ctx, int which, uint64_t* bytes);
int hdsr_debug_readback(hdsr_ctx* ctx, int which, void* host, uint64_t bytes);

typedef struct hdsr_stats {
  uint32_t refs;                  /* total tile references R */
  uint32_t small_pairs;
  uint32_t large_pairs;
  uint32_t large_splats;
  uint32_t visible_splats;
  uint32_t small_tiles;
  uint32_t large_tiles;
  uint32_t chunks;
  uint32_t segments;
  uint32_t pair_capacity;
  uint32_t overflow_reruns;       /* frames re-rendered after a capacity overflow */
  uint32_t cap_refs_tripped;      /* R > maxTotalTileRefs (gpu_tile_splat_composite.wgsl:114) */
  uint32_t cap_large_splats_tripped; /* > 65535 large splats (gpu_classify_large_splats.wgsl:52) */
  uint32_t cap_chunks_tripped;    /* > 1024 chunks (gpu_tile_bucket_sort.wgsl:126) */
  uint32_t composite_refs;        /* entries the rasterizer loaded (up to each tile's saturation) */
  uint32_t composite_max_refs;    /* the largest per-tile count of those */
  uint32_t large_tile_refs;       /* references in tiles with > 4096 entries (bucket cascade input) */
  uint32_t cell_pairs;            /* (8x8-tile cell, splat) pairs of the tile-list build */
  uint32_t sorted_refs;           /* entries the per-tile sort ordered (a frame without debug readback,
                                     G-buffer or backward sorts only the chunks its pixels reach) */
  uint32_t resumed_tiles;         /* such a frame: large tiles left unsaturated by their first chunk */
  uint32_t resumed_tile_refs;     /* their references (the bucket order their second round builds) */
  uint32_t first_chunks;          /* 1: such a frame selected the large tiles' first chunks without the
                                     bucket scatter (see HDSR_BUCKET_SCATTER) */
} hdsr_stats;
int hdsr_get_stats(hdsr_ctx* ctx, hdsr_stats* out);

/* Per-stage GPU time, from HIP events recorded on the render stream between stages (the
 * reference timestamps no compute pass; its CPU stage timing is src/splatRenderer.ts:226-283).
 * Needs HDSR_PROFILE.  Waits for outstanding renders, writes the mean milliseconds per render of
 * each stage into ms[0..HDSR_STAGE_COUNT) and the number of renders averaged into *renders, then
 * resets the accumulators. */
#define HDSR_STAGE_PROJECT 0      /* project_splats + classify_large_splats + per-splat counts */
#define HDSR_STAGE_SCAN 1         /* hierarchical scan of the per-workgroup counts */
#define HDSR_STAGE_EMIT 2         /* pair scatter (scatter_tile_refs + large scatter) */
#define HDSR_STAGE_GROUP 3        /* radix grouping on the Morton key (all passes) */
#define HDSR_STAGE_RANGES 4       /* tile offsets / counts */
#define HDSR_STAGE_CLASSIFY 5     /* classify_tiles */
#define HDSR_STAGE_BUCKET 6       /* bucket_sort (large tiles) */
#define HDSR_STAGE_SORT 7         /* tile_depth_sort + chunk_sort */
#define HDSR_STAGE_COMPOSITE 8    /* composite */
#define HDSR_STAGE_COUNT 9
int hdsr_stage_times(hdsr_ctx* ctx, float* ms, uint32_t* renders);

/* ---- Differentiable mode ------------------------------------------------------------------
 * No counterpart in the reference (its renderer has no backward; the Python preprocessing
 * package, preprocessing/splat_oracle, is the host for the autograd wrapper).  With
 * HDSR_DIFFERENTIABLE the forward is hdsr_render as usual, with opacities read as logits
 * (opacity = sigmoid(logit)).  hdsr_backward then takes dL/d(rgba) of that frame (device,
 * H*W*4 f32, same layout as the rgba output) and ADDS dL/d(parameter) into the device arrays
 * below, indexed by splat id like the inputs.  The per-tile lists and their order are held
 * fixed, as are the discrete decisions of the forward (visibility, tile boxes, the power / alpha
 * / transmittance cut-offs); the alpha clamp at 0.99 and the colour clamp to [0, 1] pass no
 * gradient.  The depth output gets no gradient.  Accumulation uses float atomics, so the low
 * bits of a gradient may differ between runs.  All pointers must be non-null device memory. */
typedef struct hdsr_grads {
  float* means;                   /* 3N */
  float* log_scales;              /* 3N */
  float* rotations;               /* 4N (w, x, y, z) */
  float* opacity_logits;          /* N */
  float* sh_data;                 /* N*(3+3K+3): dc and SH coefficients; the emissive tail stays 0 */
} hdsr_grads;
/* Waits for the last render (as hdsr_sync), then enqueues the backward on 'stream'. */
int hdsr_backward(hdsr_ctx* ctx, const float* grad_rgba, const hdsr_grads* grads, void* stream);

/* The message of the last failure on 'ctx'; with ctx NULL, of the last hdsr_create (or
   hdsr_sort_back_to_front_noalloc) that failed on the calling thread. */
const char* hdsr_last_error(hdsr_ctx* ctx);
const char* hdsr_version(void);

/* ---- Serving form -------------------------------------------------------------------------
 * No counterpart in the reference (its only frame scheduler is the browser's
 * requestAnimationFrame loop, src/main.ts:846; the closest host logic is src/renderDemand.ts).
 * A server keeps many scenes resident in device memory under a byte budget (least recently used
 * scenes are evicted; a scene is uploaded again when a request needs it), and a worker thread
 * batches the queued camera requests - any mix of scenes and resolutions - into one launch per
 * pipeline stage (up to max_views_per_batch frames), on one of 'slots' in-flight streams, so
 * requests that arrive while a batch runs are admitted into the next one as soon as a slot is
 * free.  Repeated batch shapes replay a captured hipGraph.  Each request's result is the frame
 * the plain path renders for the same scene (with the identity sortedIndices unless the scene
 * was registered with its own) and camera: batching changes no output bit. */
typedef struct hdsr_server hdsr_server;

typedef struct hdsr_server_config {
  uint64_t hbm_budget_bytes;      /* resident scene bytes; 0 = no limit */
  uint64_t max_splats_per_batch;  /* sum of scene sizes of one batch; 0 = 32M */
  uint32_t max_views_per_batch;   /* <= 64; 0 = 64 */
  uint32_t slots;                 /* batches in flight; 0 = 2 */
  uint32_t use_graphs;            /* replay captured hipGraphs for repeated batch shapes */
  uint32_t reserved;
} hdsr_server_config;

typedef struct hdsr_request {
  uint32_t scene_id;
  uint32_t width;
  uint32_t height;
  uint32_t reserved;
  const void* uniforms256;        /* the frame's 256-byte uniform block (copied at submit) */
} hdsr_request;

typedef struct hdsr_result_info {
  uint32_t done;                  /* 1 once the frame is complete */
  uint32_t width;
  uint32_t height;
  uint32_t batch_views;           /* frames in the batch this request ran in */
  double latency_ms;              /* submit -> completion observed */
  float* rgba;                    /* device H*W*4 f32 (valid until hdsr_server_release) */
  float* depth;                   /* device H*W f32 */
} hdsr_result_info;

typedef struct hdsr_server_stats {
  uint64_t frames;                /* completed requests since the last reset */
  uint64_t batches;
  uint64_t graph_replays;
  uint64_t uploads;               /* scene uploads (first admission or after eviction) */
  uint64_t evictions;
  uint64_t resident_bytes;
  uint32_t resident_scenes;
  uint32_t queued;
  double p50_ms;                  /* latency percentiles over the completed requests */
  double p99_ms;
  double mean_batch_views;
  double busy_s;                  /* wall time from the first submit to the last completion */
} hdsr_server_stats;

int hdsr_server_create(const hdsr_server_config* cfg, hdsr_server** out);
void hdsr_server_destroy(hdsr_server* srv);
/* Registers a scene from host memory (HDSR_MEM_HOST; the arrays must stay valid while the scene
 * is registered: eviction drops the device copy only).  sorted_indices may be NULL (identity). */
int hdsr_server_add_scene(hdsr_server* srv, uint32_t scene_id, const hdsr_splats* splats);
int hdsr_server_remove_scene(hdsr_server* srv, uint32_t scene_id);
int hdsr_server_submit(hdsr_server* srv, const hdsr_request* req, uint64_t* ticket);
/* hold != 0: the worker forms no new batch (batches in flight complete) until hold = 0, so a
 * burst of requests submitted meanwhile shares launches; hdsr_server_wait and hdsr_server_drain
 * release a hold.  No counterpart in the reference. */
int hdsr_server_hold(hdsr_server* srv, uint32_t hold);
/* Blocks until the ticket's frame is complete (timeout_s <= 0: no limit). */
int hdsr_server_wait(hdsr_server* srv, uint64_t ticket, double timeout_s);
int hdsr_server_result(hdsr_server* srv, uint64_t ticket, hdsr_result_info* info);
int hdsr_server_read(hdsr_server* srv, uint64_t ticket, int which, void* host, uint64_t bytes);
/* Returns the ticket's output buffers to the pool; the ticket is forgotten. */
int hdsr_server_release(hdsr_server* srv, uint64_t ticket);
/* Waits until every submitted request is complete. */
int hdsr_server_drain(hdsr_server* srv);
int hdsr_server_get_stats(hdsr_server* srv, hdsr_server_stats* out);
int hdsr_server_reset_stats(hdsr_server* srv);
const char* hdsr_server_last_error(hdsr_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* HDSR_H */
