// capi.hip - the C ABI of include/hdsr.h over one Engine (a batch of one frame).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hdsr.h"
#include "engine.h"

using namespace hdsr;

struct hdsr_ctx {
  hdsr_plan plan{};
  uint32_t flags = 0;
  int device = 0;
  Engine eng;
  // scene
  bool have_splats = false;
  bool own_splats = false;
  uint32_t n = 0, sh_degree = 0;
  uint64_t normals_len = 0;
  float *pos = nullptr, *scale = nullptr, *rot = nullptr, *opac = nullptr, *normals = nullptr, *sh = nullptr;
  uint32_t *material = nullptr, *sorted = nullptr;
  float4* packed = nullptr;       // packed records of an owned scene (the projection's input)
  bool own_sorted = false;
  // frame
  bool have_frame = false;
  uint32_t frame_sh_degree = 0;   // word 25 of the last uniforms
  FrameDesc frame{};
  // outputs
  float* rgba = nullptr;
  float* depth = nullptr;
  float* user_rgba = nullptr;     // caller-bound outputs (hdsr_bind_outputs), or null
  float* user_depth = nullptr;
  uint32_t* gbn = nullptr;
  uint32_t* gbm = nullptr;
  uint32_t* gbn_filtered = nullptr;  // screen-space normals output (allocated on first use)
  uint16_t* rgba16 = nullptr;
  float* t_final = nullptr;       // differentiable mode
  uint32_t* last_idx = nullptr;
  uint2* ds_work = nullptr;       // depth-sort scratch (hdsr_sort_back_to_front), on first use
  uint16_t* lit = nullptr;        // deferred lighting output (hdsr_deferred_lighting), on first use
  float* lut_dev = nullptr;       // the three material-curve tables on the device (3 x 256)
  std::vector<float> lut_host;    // their last uploaded values
  hipStream_t last_stream = nullptr;
  bool rendered = false;
  bool settled = false;  // the last render's capacity check has run (hdsr_sync since the render)
  uint32_t reruns = 0;
  std::string err;
};

namespace {

int fail(hdsr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPX(c, expr)                                                              \
  do {                                                                             \
    const hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) return fail(c, HDSR_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void free_scene(hdsr_ctx* c) {
  if (c->own_splats) {
    void* p[] = {c->pos, c->scale, c->rot, c->opac, c->normals, c->sh, c->material, c->packed};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
  c->packed = nullptr;
  if (c->own_sorted && c->sorted) (void)hipFree(c->sorted);
  c->pos = c->scale = c->rot = c->opac = c->normals = c->sh = nullptr;
  c->material = c->sorted = nullptr;
  c->have_splats = c->own_splats = c->own_sorted = false;
}

template <typename T>
int dev_copy(hdsr_ctx* c, T** dst, const T* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return 0;
  HIPX(c, hipMalloc((void**)dst, count * sizeof(T)));
  HIPX(c, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

uint32_t sh_stride(uint32_t deg) { return sh_row_floats(deg); }

// The frame's uniform shDegree against the bound shData (the shader derives the row stride from
// the uniform, gpu_project_splats.wgsl:294-296, so a mismatch would read other splats' rows).
int check_frame_degree(hdsr_ctx* c) {
  if (c->have_splats && c->have_frame && c->frame_sh_degree != c->sh_degree)
    return fail(c, HDSR_E_INVALID_ARG,
                "uniform shDegree (" + std::to_string(c->frame_sh_degree) + ") differs from the bound shData's degree (" +
                    std::to_string(c->sh_degree) + ")");
  return HDSR_OK;
}

void fill_frame_scene(hdsr_ctx* c) {
  FrameDesc& F = c->frame;
  F.pos = c->pos;
  F.logscale = c->scale;
  F.rot = c->rot;
  F.opacity = c->opac;
  F.sorted = c->sorted;
  F.material = c->material;
  F.normals = c->normals;
  F.normals_len = (uint32_t)std::min<uint64_t>(c->normals_len, 0xFFFFFFFFull);
  F.shdata = c->sh;
  F.n = c->n;
  F.sh_degree = c->sh_degree;
  F.sh_stride = sh_stride(c->sh_degree);
  F.packed = c->packed;
  F.pk_stride4 = packed_stride4(F.sh_stride);
  F.width = c->plan.viewport_width;
  F.height = c->plan.viewport_height;
  F.rgba = c->user_rgba ? c->user_rgba : c->rgba;
  F.out_depth = c->user_depth ? c->user_depth : c->depth;
  F.gb_normal = c->gbn;
  F.gb_material = c->gbm;
  F.rgba16 = c->rgba16;
  F.diff = (c->flags & HDSR_DIFFERENTIABLE) ? 1u : 0u;
  F.t_final = c->t_final;
  F.last_idx = c->last_idx;
}

}  // namespace

extern "C" {

const char* hdsr_version(void) { return "hdsr-mi355x 0.1 (gfx950)"; }

// Failures before a context exists (hdsr_create) leave their message here, per host thread.
static thread_local std::string g_create_err = "null context";

static int create_fail(int code, const char* what, hipError_t e) {
  g_create_err = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  (void)hipGetLastError();
  return code;
}

const char* hdsr_last_error(hdsr_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int hdsr_plan_compute(const hdsr_plan_in* in, hdsr_plan* out) {
  if (!in || !out) return HDSR_E_INVALID_ARG;
  if (in->tile_size_px == 0 || in->viewport_width == 0 || in->viewport_height == 0) return HDSR_E_INVALID_ARG;
  hdsr_plan p{};
  p.viewport_width = in->viewport_width;
  p.viewport_height = in->viewport_height;
  p.tile_size_px = in->tile_size_px;
  p.tile_columns = (in->viewport_width + in->tile_size_px - 1) / in->tile_size_px;
  p.tile_rows = (in->viewport_height + in->tile_size_px - 1) / in->tile_size_px;
  p.tile_count = p.tile_columns * p.tile_rows;
  p.morton_tile_count = hdsr_morton2(p.tile_columns - 1, p.tile_rows - 1) + 1;
  p.splat_count = in->splat_count;
  const double mult = in->tile_entry_multiplier > 0.0f ? (double)in->tile_entry_multiplier : 16.0;
  const double want = std::ceil((double)in->splat_count * mult);
  p.max_total_tile_refs = (uint32_t)std::min(want, 65535.0 * 256.0);
  *out = p;
  return HDSR_OK;
}

int hdsr_create(const hdsr_plan* plan, uint32_t flags, hdsr_ctx** out) {
  if (!plan || !out) return HDSR_E_INVALID_ARG;
  *out = nullptr;
  if ((flags & HDSR_DIFFERENTIABLE) && (flags & (HDSR_GBUFFER | HDSR_F16_COMPOSITE | HDSR_REF_CAPS)))
    return HDSR_E_UNSUPPORTED;
  if (plan->tile_size_px != HDSR_TILE_PX) return HDSR_E_INVALID_ARG;
  if (plan->tile_columns == 0 || plan->tile_rows == 0 || plan->tile_columns > 255 || plan->tile_rows > 255)
    return HDSR_E_INVALID_ARG;  // u8 tile-box packing (gpu_project_splats.wgsl:229-231)
  hdsr_ctx* c = new hdsr_ctx();
  c->plan = *plan;
  c->flags = flags;
  if (hipError_t e = hipGetDevice(&c->device); e != hipSuccess) {
    delete c;
    return create_fail(HDSR_E_HIP, "hipGetDevice", e);
  }
  const size_t px = (size_t)plan->viewport_width * plan->viewport_height;
  // every output the flags ask for; on a failure the ones already allocated are freed
  bool ok = hipMalloc((void**)&c->rgba, px * 16) == hipSuccess && hipMalloc((void**)&c->depth, px * 4) == hipSuccess;
  if (ok && (flags & HDSR_GBUFFER))
    ok = hipMalloc((void**)&c->gbn, px * 4) == hipSuccess && hipMalloc((void**)&c->gbm, px * 16) == hipSuccess;
  if (ok && (flags & HDSR_OUT_RGBA16)) ok = hipMalloc((void**)&c->rgba16, px * 8) == hipSuccess;
  if (ok && (flags & HDSR_DIFFERENTIABLE))
    ok = hipMalloc((void**)&c->t_final, px * 4) == hipSuccess && hipMalloc((void**)&c->last_idx, px * 4) == hipSuccess;
  if (!ok) {
    const hipError_t e = hipGetLastError();
    hdsr_destroy(c);
    return create_fail(HDSR_E_OOM, "hipMalloc of the outputs", e);
  }
  if (flags & HDSR_DIFFERENTIABLE) c->eng.set_diff(true);
  c->eng.set_f16((flags & HDSR_F16_COMPOSITE) != 0);
  const uint64_t base_cap = std::max<uint64_t>((uint64_t)plan->splat_count * 16u, 1u << 16);
  c->eng.set_min_capacity(base_cap);
  if (flags & HDSR_REF_CAPS) c->eng.set_ref_caps(plan->max_total_tile_refs);
  c->eng.set_profile((flags & HDSR_PROFILE) != 0);
  c->eng.set_bucket_scatter((flags & HDSR_BUCKET_SCATTER) != 0);
  *out = c;
  return HDSR_OK;
}

void hdsr_destroy(hdsr_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  free_scene(c);
  void* p[] = {c->rgba, c->depth, c->gbn, c->gbm, c->rgba16, c->t_final, c->last_idx, c->ds_work, c->gbn_filtered,
               c->lit, c->lut_dev};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete c;
}

int hdsr_bind_splats(hdsr_ctx* c, const hdsr_splats* s) {
  if (!c || !s) return HDSR_E_INVALID_ARG;
  if (s->count != c->plan.splat_count) return fail(c, HDSR_E_INVALID_ARG, "splat count differs from the plan");
  uint32_t deg = 0;
  const char* why = "";
  if (!resolve_sh_degree(s->count, s->sh_degree, s->sh_data_len, &deg, &why)) return fail(c, HDSR_E_INVALID_ARG, why);
  if (s->count && (!s->positions || !s->scales || !s->rotations || !s->opacities || !s->materials || !s->sh_data ||
                   !s->sorted_indices))
    return fail(c, HDSR_E_INVALID_ARG, "missing splat buffer");
  if (s->normals_len && !s->normals) return fail(c, HDSR_E_INVALID_ARG, "normals_len without normals");
  free_scene(c);
  const size_t n = s->count;
  c->n = s->count;
  c->sh_degree = deg;
  c->normals_len = s->normals_len;
  if (s->memory == HDSR_MEM_DEVICE) {
    c->pos = const_cast<float*>(s->positions);
    c->scale = const_cast<float*>(s->scales);
    c->rot = const_cast<float*>(s->rotations);
    c->opac = const_cast<float*>(s->opacities);
    c->material = const_cast<uint32_t*>(s->materials);
    c->normals = const_cast<float*>(s->normals);
    c->sh = const_cast<float*>(s->sh_data);
    c->sorted = const_cast<uint32_t*>(s->sorted_indices);
    c->own_splats = false;
    c->own_sorted = false;
  } else if (s->memory == HDSR_MEM_HOST) {
    c->own_splats = true;
    c->own_sorted = true;
    int rc;
    if ((rc = dev_copy(c, &c->pos, s->positions, n * 3))) return rc;
    if ((rc = dev_copy(c, &c->scale, s->scales, n * 3))) return rc;
    if ((rc = dev_copy(c, &c->rot, s->rotations, n * 4))) return rc;
    if ((rc = dev_copy(c, &c->opac, s->opacities, n))) return rc;
    if ((rc = dev_copy(c, &c->material, s->materials, n))) return rc;
    if ((rc = dev_copy(c, &c->normals, s->normals, (size_t)s->normals_len))) return rc;
    if ((rc = dev_copy(c, &c->sh, s->sh_data, n * sh_stride(deg)))) return rc;
    if ((rc = dev_copy(c, &c->sorted, s->sorted_indices, n))) return rc;
    if (n) {
      // one 128-byte-aligned record per splat for the projection's gathers
      const uint32_t stride = sh_stride(deg);
      HIPX(c, hipMalloc((void**)&c->packed, n * (size_t)packed_stride4(stride) * sizeof(float4)));
      launch_pack_scene(c->pos, c->scale, c->rot, c->opac, c->material, c->sh, (uint32_t)n, stride, c->packed,
                        nullptr);
      HIPX(c, hipGetLastError());
      HIPX(c, hipDeviceSynchronize());
    }
  } else {
    return fail(c, HDSR_E_INVALID_ARG, "memory must be HDSR_MEM_HOST or HDSR_MEM_DEVICE");
  }
  c->have_splats = true;
  return check_frame_degree(c);
}

int hdsr_update_sorted_indices(hdsr_ctx* c, const uint32_t* sorted, uint32_t memory) {
  if (!c || (!sorted && c->n)) return HDSR_E_INVALID_ARG;
  if (!c->have_splats) return fail(c, HDSR_E_STATE, "bind splats first");
  if (memory == HDSR_MEM_DEVICE) {
    if (c->own_sorted && c->sorted) (void)hipFree(c->sorted);
    c->sorted = const_cast<uint32_t*>(sorted);
    c->own_sorted = false;
    return HDSR_OK;
  }
  if (!c->own_sorted) {
    c->sorted = nullptr;
    if (c->n) HIPX(c, hipMalloc((void**)&c->sorted, (size_t)c->n * 4));
    c->own_sorted = true;
  }
  if (c->n) {
    HIPX(c, hipDeviceSynchronize());
    HIPX(c, hipMemcpy(c->sorted, sorted, (size_t)c->n * 4, hipMemcpyHostToDevice));
  }
  return HDSR_OK;
}

int hdsr_set_frame(hdsr_ctx* c, const void* u) {
  if (!c || !u) return HDSR_E_INVALID_ARG;
  FrameDesc f{};
  parse_uniforms(u, &f);
  const uint32_t* w = static_cast<const uint32_t*>(u);
  if (f.tiles_x != c->plan.tile_columns || f.tiles_y != c->plan.tile_rows)
    return fail(c, HDSR_E_INVALID_ARG, "uniform tile grid differs from the plan");
  if (w[22] != c->plan.splat_count) return fail(c, HDSR_E_INVALID_ARG, "uniform splatCount differs from the plan");
  if (w[25] > 3) return fail(c, HDSR_E_INVALID_ARG, "uniform shDegree must be 0..3");
  if (c->have_splats && w[25] != c->sh_degree)
    return fail(c, HDSR_E_INVALID_ARG,
                "uniform shDegree (" + std::to_string(w[25]) + ") differs from the bound shData's degree (" +
                    std::to_string(c->sh_degree) + ")");
  c->frame = f;
  c->frame_sh_degree = w[25];
  c->have_frame = true;
  return HDSR_OK;
}

int hdsr_sort_back_to_front(hdsr_ctx* c, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  if (!c->have_splats || !c->have_frame) return fail(c, HDSR_E_STATE, "bind splats and set a frame first");
  const uint32_t n = c->n;
  if (!n) return HDSR_OK;
  if (!c->ds_work) HIPX(c, hipMalloc((void**)&c->ds_work, (size_t)depth_sort_padded(n) * sizeof(uint2)));
  if (!c->own_sorted) {  // a borrowed buffer is not written: switch to an owned one
    uint32_t* own = nullptr;
    HIPX(c, hipMalloc((void**)&own, (size_t)n * 4));
    c->sorted = own;
    c->own_sorted = true;
  }
  launch_depth_sort(c->pos, n, c->frame.view, c->sorted, c->ds_work, static_cast<hipStream_t>(stream));
  HIPX(c, hipGetLastError());
  return HDSR_OK;
}

// ---- allocation-free depth order -----------------------------------------------------------
// Scratch of hdsr_sort_back_to_front_noalloc: the padded (key, index) pairs of the bitonic network,
// HDSR_NOALLOC_MAX_SPLATS of them (32 MiB), in a file-scope device array that every GPU the
// library runs on holds until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and only that address is handed to kernels.
// k_depth_keys writes every pair of the padded range before the network reads it, so nothing in
// it has to start at zero.  All calls on one GPU share it: calls that overlap on one GPU race.
static __device__ __align__(256) uint2 scratch[HDSR_NOALLOC_MAX_SPLATS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on 'device' (current on this thread), looked up on first use.
hipError_t scratch_on(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

int hdsr_sort_back_to_front_noalloc(const float* positions, uint32_t count, const void* uniforms256,
                                    uint32_t* sorted_out, void* stream) {
  if (!uniforms256 || (count && (!positions || !sorted_out))) {
    g_create_err = "hdsr_sort_back_to_front_noalloc: null argument";
    return HDSR_E_INVALID_ARG;
  }
  if (count > HDSR_NOALLOC_MAX_SPLATS) {
    g_create_err = "hdsr_sort_back_to_front_noalloc: " + std::to_string(count) + " splats exceed the scratch of " +
                   std::to_string(HDSR_NOALLOC_MAX_SPLATS);
    return HDSR_E_CAPACITY;
  }
  if (!count) return HDSR_OK;
  int device = -1;
  hipError_t e = hipGetDevice(&device);
  if (e == hipSuccess && (device < 0 || device >= kMaxDevices)) e = hipErrorInvalidDevice;
  void* work = nullptr;
  if (e == hipSuccess) e = scratch_on(device, &work);
  if (e != hipSuccess) return create_fail(HDSR_E_HIP, "hdsr_sort_back_to_front_noalloc: scratch lookup", e);
  FrameDesc f{};
  parse_uniforms(uniforms256, &f);
  launch_depth_sort(positions, count, f.view, sorted_out, static_cast<uint2*>(work), static_cast<hipStream_t>(stream));
  e = hipGetLastError();
  if (e != hipSuccess) return create_fail(HDSR_E_HIP, "hdsr_sort_back_to_front_noalloc: launch", e);
  return HDSR_OK;
}

// The context's device current on the calling thread (a context may be driven from another host
// thread than the one that created it, one thread at a time, e.g. the N-API addon's worker).
static int use_device(hdsr_ctx* c) {
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == c->device) return HDSR_OK;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, HDSR_E_HIP, "hipSetDevice failed");
  return HDSR_OK;
}

int hdsr_render(hdsr_ctx* c, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  if (!c->have_splats || !c->have_frame) return fail(c, HDSR_E_STATE, "bind splats and set a frame first");
  if (int rc = check_frame_degree(c)) return rc;
  if (int rc = use_device(c)) return rc;
  fill_frame_scene(c);
  const int rc = c->eng.run(&c->frame, 1, static_cast<hipStream_t>(stream), (c->flags & HDSR_GBUFFER) != 0,
                            (c->flags & HDSR_DEBUG) != 0);
  if (rc) return fail(c, rc, c->eng.err);
  c->last_stream = static_cast<hipStream_t>(stream);
  c->rendered = true;
  c->settled = false;
  return HDSR_OK;
}

int hdsr_composite_only(hdsr_ctx* c, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  if (!c->rendered) return fail(c, HDSR_E_STATE, "render a frame first");
  fill_frame_scene(c);
  const int rc = c->eng.composite_only(&c->frame, 1, static_cast<hipStream_t>(stream), (c->flags & HDSR_GBUFFER) != 0);
  if (rc) return fail(c, rc, c->eng.err);
  return HDSR_OK;
}

int hdsr_sync(hdsr_ctx* c) {
  if (!c) return HDSR_E_INVALID_ARG;
  if (!c->rendered) return HDSR_OK;
  if (int rc = use_device(c)) return rc;
  for (int attempt = 0; attempt < 4; ++attempt) {
    int rc = c->eng.wait();
    if (rc) return fail(c, rc, c->eng.err);
    HIPX(c, hipStreamSynchronize(c->last_stream));
    if (!c->eng.overflowed()) {
      c->settled = true;
      return HDSR_OK;
    }
    c->eng.grow_for_overflow();
    c->reruns++;
    rc = c->eng.run(&c->frame, 1, c->last_stream, (c->flags & HDSR_GBUFFER) != 0, (c->flags & HDSR_DEBUG) != 0);
    if (rc) return fail(c, rc, c->eng.err);
  }
  return fail(c, HDSR_E_CAPACITY, "pair capacity still overflowing after regrowth");
}

int hdsr_bind_outputs(hdsr_ctx* c, float* rgba, float* depth) {
  if (!c) return HDSR_E_INVALID_ARG;
  c->user_rgba = rgba;
  c->user_depth = depth;
  return HDSR_OK;
}

int hdsr_get_outputs(hdsr_ctx* c, hdsr_outputs* o) {
  if (!c || !o) return HDSR_E_INVALID_ARG;
  o->rgba = c->user_rgba ? c->user_rgba : c->rgba;
  o->depth = c->user_depth ? c->user_depth : c->depth;
  o->gbuffer_normal = c->gbn;
  o->gbuffer_material = c->gbm;
  o->rgba16 = c->rgba16;
  o->width = c->plan.viewport_width;
  o->height = c->plan.viewport_height;
  o->lit = c->lit;
  return HDSR_OK;
}

int hdsr_readback(hdsr_ctx* c, int which, void* host, uint64_t bytes) {
  if (!c || !host) return HDSR_E_INVALID_ARG;
  int rc = hdsr_sync(c);
  if (rc) return rc;
  const size_t px = (size_t)c->plan.viewport_width * c->plan.viewport_height;
  const void* src = nullptr;
  size_t need = 0;
  switch (which) {
    case HDSR_OUT_RGBA: src = c->user_rgba ? c->user_rgba : c->rgba; need = px * 16; break;
    case HDSR_OUT_DEPTH: src = c->user_depth ? c->user_depth : c->depth; need = px * 4; break;
    case HDSR_OUT_GB_NORMAL: src = c->gbn; need = px * 4; break;
    case HDSR_OUT_GB_MATERIAL: src = c->gbm; need = px * 16; break;
    case HDSR_OUT_RGBA16_BUF: src = c->rgba16; need = px * 8; break;
    case HDSR_OUT_T_FINAL: src = c->t_final; need = px * 4; break;
    case HDSR_OUT_LAST_INDEX: src = c->last_idx; need = px * 4; break;
    case HDSR_OUT_FILTERED_NORMAL: src = c->gbn_filtered; need = px * 4; break;
    case HDSR_OUT_LIT: src = c->lit; need = px * 8; break;
    default: return fail(c, HDSR_E_INVALID_ARG, "unknown output");
  }
  if (!src) return fail(c, HDSR_E_STATE, "output not enabled by the create flags");
  if (bytes < need) return fail(c, HDSR_E_INVALID_ARG, "host buffer too small");
  HIPX(c, hipMemcpy(host, src, need, hipMemcpyDeviceToHost));
  return HDSR_OK;
}

void hdsr_normal_params_default(hdsr_normal_params* p) {
  if (!p) return;
  p->recovery_mode = 0;
  p->grazing_start = 0.45f;
  p->grazing_end = 0.80f;
  p->spatial_sigma = 3.0f;
  p->depth_sigma = 0.05f;
  p->normal_sigma = 1.5f;
  p->kernel_radius = 3;
}

namespace {

// mat4Inverse (src/splatRenderer.ts:148-168): cofactors in double from the fp32 matrix, stored to
// fp32; false when |det| < 1e-10 (the reference then skips the deferred passes).
bool mat4_inverse_ref(const float* m32, float* o) {
  double m[16];
  for (int i = 0; i < 16; ++i) m[i] = (double)m32[i];
  const double a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3], a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
  const double a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11], a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
  const double b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
  const double b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
  const double b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
  const double b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
  double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  if (std::fabs(det) < 1e-10) return false;
  det = 1.0 / det;
  o[0] = (float)((a11 * b11 - a12 * b10 + a13 * b09) * det);
  o[1] = (float)((a02 * b10 - a01 * b11 - a03 * b09) * det);
  o[2] = (float)((a31 * b05 - a32 * b04 + a33 * b03) * det);
  o[3] = (float)((a22 * b04 - a21 * b05 - a23 * b03) * det);
  o[4] = (float)((a12 * b08 - a10 * b11 - a13 * b07) * det);
  o[5] = (float)((a00 * b11 - a02 * b08 + a03 * b07) * det);
  o[6] = (float)((a32 * b02 - a30 * b05 - a33 * b01) * det);
  o[7] = (float)((a20 * b05 - a22 * b02 + a23 * b01) * det);
  o[8] = (float)((a10 * b10 - a11 * b08 + a13 * b06) * det);
  o[9] = (float)((a01 * b08 - a00 * b10 - a03 * b06) * det);
  o[10] = (float)((a30 * b04 - a31 * b02 + a33 * b00) * det);
  o[11] = (float)((a21 * b02 - a20 * b04 - a23 * b00) * det);
  o[12] = (float)((a11 * b07 - a10 * b09 - a12 * b06) * det);
  o[13] = (float)((a00 * b09 - a01 * b07 + a02 * b06) * det);
  o[14] = (float)((a31 * b01 - a30 * b03 - a32 * b00) * det);
  o[15] = (float)((a20 * b03 - a21 * b01 + a22 * b00) * det);
  return true;
}

int deferred_ready(hdsr_ctx* c) {
  if (!(c->flags & HDSR_GBUFFER)) return fail(c, HDSR_E_STATE, "create the context with HDSR_GBUFFER");
  if (!c->rendered) return fail(c, HDSR_E_STATE, "render a frame first");
  // the render is final only once a capacity overflow has been re-run
  if (!c->settled) {
    const int rc = hdsr_sync(c);
    if (rc) return rc;
  }
  if (!c->gbn_filtered) {
    const size_t px = (size_t)c->plan.viewport_width * c->plan.viewport_height;
    HIPX(c, hipMalloc((void**)&c->gbn_filtered, px * 4));
  }
  return HDSR_OK;
}

}  // namespace

int hdsr_screen_space_normals(hdsr_ctx* c, const hdsr_normal_params* params, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  int rc = deferred_ready(c);
  if (rc) return rc;
  hdsr_normal_params p;
  hdsr_normal_params_default(&p);
  if (params) p = *params;
  if (p.recovery_mode > 2) return fail(c, HDSR_E_INVALID_ARG, "recovery_mode must be 0, 1 or 2");
  float inv[16];
  if (!mat4_inverse_ref(c->frame.vp, inv)) return fail(c, HDSR_E_INVALID_ARG, "viewProj is singular");
  const float* depth = c->user_depth ? c->user_depth : c->depth;
  launch_screen_space_normals(depth, c->gbn, c->gbn_filtered, c->plan.viewport_width, c->plan.viewport_height, inv,
                              c->frame.cam, p.recovery_mode, p.grazing_start, p.grazing_end,
                              static_cast<hipStream_t>(stream));
  HIPX(c, hipGetLastError());
  c->last_stream = static_cast<hipStream_t>(stream);
  return HDSR_OK;
}

int hdsr_bilateral_normal_filter(hdsr_ctx* c, const hdsr_normal_params* params, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  int rc = deferred_ready(c);
  if (rc) return rc;
  hdsr_normal_params p;
  hdsr_normal_params_default(&p);
  if (params) p = *params;
  const float* depth = c->user_depth ? c->user_depth : c->depth;
  if (launch_bilateral_normals(c->gbn_filtered, depth, c->gbn, c->plan.viewport_width, c->plan.viewport_height,
                               p.spatial_sigma, p.depth_sigma, p.normal_sigma, (int)p.kernel_radius,
                               static_cast<hipStream_t>(stream)) != 0)
    return fail(c, HDSR_E_INVALID_ARG, "kernel_radius must be at most 8");
  HIPX(c, hipGetLastError());
  c->last_stream = static_cast<hipStream_t>(stream);
  return HDSR_OK;
}

void hdsr_lighting_params_default(hdsr_lighting_params* p) {
  if (!p) return;
  *p = hdsr_lighting_params{};
  // the reference viewer's fixed light (src/main.ts:145-149, 273-279): azimuth 0.8, elevation 0.6
  const double ce = std::cos(0.6);
  p->light_direction[0] = std::cos(0.8) * ce;
  p->light_direction[1] = -std::sin(0.6);
  p->light_direction[2] = std::sin(0.8) * ce;
  p->light_intensity = 3.0;
  p->ambient_intensity = 0.12;
  p->specular_only = 0;
  p->emissive_intensity = 3.0;
  p->emissive_threshold = 0.05;
  p->env_intensity = 1.0;
  p->env_rotation = 0.0;
  p->exposure = 1.0;
  p->roughness_curve = p->metalness_curve = p->albedo_curve = hdsr_material_curve{1.0, 0.0, 1.0};
}

void hdsr_write_lighting_params(const hdsr_lighting_params* p, const float* inv, const float* cam, uint32_t w,
                                uint32_t h, float* o) {
  if (!o) return;
  hdsr_lighting_params d;
  hdsr_lighting_params_default(&d);
  if (!p) p = &d;
  std::memset(o, 0, 44 * sizeof(float));
  o[0] = (float)w;
  o[1] = (float)h;
  o[2] = 0.65f;  // fallback roughness / metallic (the shader reads the G-buffer's)
  o[3] = 0.0f;
  for (int i = 0; i < 16; ++i) o[4 + i] = inv ? inv[i] : 0.0f;
  for (int i = 0; i < 3; ++i) o[20 + i] = cam ? cam[i] : 0.0f;
  for (int i = 0; i < 3; ++i) o[24 + i] = (float)p->light_direction[i];
  o[28] = 1.0f;  // warm white
  o[29] = (float)0.95;
  o[30] = (float)0.9;
  o[31] = (float)p->light_intensity;
  o[32] = (float)p->ambient_intensity;
  o[33] = (float)p->ambient_intensity;
  o[34] = (float)(p->ambient_intensity * 1.1);
  o[35] = p->specular_only ? 1.0f : 0.0f;
  o[36] = (float)p->emissive_intensity;
  o[37] = (float)p->emissive_threshold;
  o[38] = (float)p->env_intensity;
  o[39] = (float)p->env_rotation;
  o[40] = (float)p->exposure;
}

void hdsr_material_curve_lut(const hdsr_material_curve* curve, float* out) {
  if (!out) return;
  const hdsr_material_curve id{1.0, 0.0, 1.0};
  const hdsr_material_curve& c = curve ? *curve : id;
  // JavaScript's Math.min / Math.max propagate NaN
  auto clamp01 = [](double v) { return v != v ? v : (v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v)); };
  for (int i = 0; i < 256; ++i) {
    double v = i / 255.0;
    v = (v - 0.5) * c.contrast + 0.5;
    v += c.brightness;
    v = clamp01(v);
    v = std::pow(v, c.gamma);
    const float f = (float)clamp01(v);  // Float32Array store
    uint32_t bits;
    std::memcpy(&bits, &f, 4);
    // f32ToF16 of uploadLUT: mantissa truncated, subnormal results flushed to signed zero
    const uint32_t sign = (bits >> 16) & 0x8000u;
    const int e = (int)((bits >> 23) & 0xFFu) - 127 + 15;
    const uint32_t frac = bits & 0x7FFFFFu;
    uint32_t h16;
    if (e <= 0) h16 = sign;
    else if (e >= 31) h16 = sign | 0x7C00u;
    else h16 = sign | ((uint32_t)e << 10) | (frac >> 13);
    out[i] = hdsr_f16_to_f32(h16);
  }
}

int hdsr_deferred_lighting(hdsr_ctx* c, const hdsr_lighting_params* params, void* stream) {
  if (!c) return HDSR_E_INVALID_ARG;
  int rc = deferred_ready(c);
  if (rc) return rc;
  hdsr_lighting_params p;
  hdsr_lighting_params_default(&p);
  if (params) p = *params;
  float inv[16];
  const float* vp = p.lighting_view_proj ? p.lighting_view_proj : c->frame.vp;
  if (!mat4_inverse_ref(vp, inv)) return fail(c, HDSR_E_INVALID_ARG, "lighting viewProj is singular");
  const float* cam = p.lighting_camera_position ? p.lighting_camera_position : c->frame.cam;
  LightingBlock blk;
  hdsr_write_lighting_params(&p, inv, cam, c->plan.viewport_width, c->plan.viewport_height, blk.f);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t px = (size_t)c->plan.viewport_width * c->plan.viewport_height;
  if (!c->lit) HIPX(c, hipMalloc((void**)&c->lit, std::max<size_t>(px, 1) * 8));
  if (!c->lut_dev) HIPX(c, hipMalloc((void**)&c->lut_dev, 768 * sizeof(float)));
  std::vector<float> lut(768);
  hdsr_material_curve_lut(&p.roughness_curve, lut.data());
  hdsr_material_curve_lut(&p.metalness_curve, lut.data() + 256);
  hdsr_material_curve_lut(&p.albedo_curve, lut.data() + 512);
  if (lut != c->lut_host) {  // materialCurves.update: upload only what changed
    HIPX(c, hipMemcpyAsync(c->lut_dev, lut.data(), 768 * sizeof(float), hipMemcpyHostToDevice, s));
    HIPX(c, hipStreamSynchronize(s));  // the host table is a temporary
    c->lut_host = lut;
  }
  const float* rgba = c->user_rgba ? c->user_rgba : c->rgba;
  const float* depth = c->user_depth ? c->user_depth : c->depth;
  launch_deferred_lighting(blk, rgba, depth, c->gbn, c->gbm, p.ao, p.bloom, c->lut_dev, c->lit,
                           c->plan.viewport_width, c->plan.viewport_height, s);
  HIPX(c, hipGetLastError());
  c->last_stream = s;
  return HDSR_OK;
}

int hdsr_backward(hdsr_ctx* c, const float* grad_rgba, const hdsr_grads* g, void* stream) {
  if (!c || !grad_rgba || !g) return HDSR_E_INVALID_ARG;
  if (!(c->flags & HDSR_DIFFERENTIABLE)) return fail(c, HDSR_E_STATE, "create the context with HDSR_DIFFERENTIABLE");
  if (!c->rendered) return fail(c, HDSR_E_STATE, "render a frame first");
  if (c->n && (!g->means || !g->log_scales || !g->rotations || !g->opacity_logits || !g->sh_data))
    return fail(c, HDSR_E_INVALID_ARG, "missing gradient buffer");
  int rc = hdsr_sync(c);
  if (rc) return rc;
  DiffGrads d{g->means, g->log_scales, g->rotations, g->opacity_logits, g->sh_data};
  rc = c->eng.backward(c->n, grad_rgba, d, static_cast<hipStream_t>(stream));
  if (rc) return fail(c, rc, c->eng.err);
  return HDSR_OK;
}

int hdsr_get_stats(hdsr_ctx* c, hdsr_stats* o) {
  if (!c || !o) return HDSR_E_INVALID_ARG;
  int rc = hdsr_sync(c);
  if (rc) return rc;
  memset(o, 0, sizeof(*o));
  const DevStats& s = c->eng.host_stats();
  // the frame's references (the device count is clamped to the pair capacity, which only the
  // reference's caps leave below them once the frame is settled)
  const uint64_t refs = (uint64_t)s.small_pairs + s.large_pairs;
  o->refs = refs > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)refs;
  o->small_pairs = s.small_pairs;
  o->large_pairs = s.large_pairs;
  o->large_splats = s.large_splats;
  o->visible_splats = s.visible;
  o->small_tiles = s.small_tiles;
  o->large_tiles = s.large_tiles;
  o->chunks = s.chunks;
  o->segments = s.segments;
  o->pair_capacity = c->eng.bufs().cap;
  o->overflow_reruns = c->reruns;
  o->cap_refs_tripped = refs > c->plan.max_total_tile_refs;
  o->cap_large_splats_tripped = s.large_splats > 65535u;
  o->cap_chunks_tripped = s.chunks > 1024u;
  o->composite_refs = s.composite_refs;
  o->composite_max_refs = s.composite_max_refs;
  o->large_tile_refs = s.scratch_used;
  o->cell_pairs = s.cell_pairs;
  // a deferred frame sorts the small tiles, the first 4096-entry chunk of every large tile and the
  // other chunks of the tiles resumed; a one-pass frame sorts every entry
  const bool deferred = c->eng.last_frame_deferred();
  o->sorted_refs = deferred ? (s.refs - s.scratch_used) + HDSR_MAX_TILE_ENTRIES * s.large_tiles + s.resumed_refs
                            : s.refs;
  o->resumed_tiles = deferred ? s.pend_tiles : 0u;
  o->resumed_tile_refs = deferred ? s.resumed_refs + HDSR_MAX_TILE_ENTRIES * s.pend_tiles : 0u;
  o->first_chunks = deferred && c->eng.last_frame_first_chunks() ? 1u : 0u;
  return HDSR_OK;
}

int hdsr_stage_times(hdsr_ctx* c, float* ms, uint32_t* renders) {
  if (!c || !ms || !renders) return HDSR_E_INVALID_ARG;
  if (!(c->flags & HDSR_PROFILE)) return fail(c, HDSR_E_STATE, "create the context with HDSR_PROFILE");
  return c->eng.stage_times(ms, renders);
}

int hdsr_debug_size(hdsr_ctx* c, int which, uint64_t* bytes) {
  if (!c || !bytes) return HDSR_E_INVALID_ARG;
  int rc = hdsr_sync(c);
  if (rc) return rc;
  if (which == HDSR_DBG_SORTED) {
    if (!c->have_splats) return fail(c, HDSR_E_STATE, "bind splats first");
    *bytes = (uint64_t)c->n * 4;
    return HDSR_OK;
  }
  if (!c->rendered) return fail(c, HDSR_E_STATE, "render a frame first");
  const DevStats& s = c->eng.host_stats();
  const uint64_t M = c->plan.morton_tile_count;
  // the entry arrays hold at most the pair capacity (below the references with HDSR_REF_CAPS)
  const uint64_t R = std::min<uint64_t>(s.refs, c->eng.bufs().cap);
  switch (which) {
    case HDSR_DBG_PROJCACHE:
      // the G-buffer-only words of the record are written by renders that need them
      if (!(c->flags & (HDSR_DEBUG | HDSR_GBUFFER)))
        return fail(c, HDSR_E_STATE, "create the context with HDSR_DEBUG (or HDSR_GBUFFER) to read the projection cache");
      *bytes = (uint64_t)c->n * 56;
      break;
    case HDSR_DBG_DEPTH: *bytes = (uint64_t)c->n * 4; break;
    case HDSR_DBG_TILE_COUNTS:
    case HDSR_DBG_TILE_OFFSETS: *bytes = M * 4; break;
    case HDSR_DBG_ENTRIES: *bytes = R * 4; break;
    case HDSR_DBG_ENTRIES_CANON:
    case HDSR_DBG_KEYS20:
      if (!(c->flags & HDSR_DEBUG)) return fail(c, HDSR_E_STATE, "create the context with HDSR_DEBUG");
      *bytes = R * 4;
      break;
    case HDSR_DBG_SMALL_LIST: *bytes = (uint64_t)s.small_tiles * 4; break;
    case HDSR_DBG_LARGE_LIST:
    case HDSR_DBG_LARGE_BASES: *bytes = (uint64_t)s.large_tiles * 4; break;
    case HDSR_DBG_CHUNKS: *bytes = (uint64_t)s.chunks * 8; break;
    case HDSR_DBG_LARGE_SPLATS: *bytes = (uint64_t)s.large_splats * 4; break;
    default: return fail(c, HDSR_E_INVALID_ARG, "unknown debug buffer");
  }
  return HDSR_OK;
}

int hdsr_debug_readback(hdsr_ctx* c, int which, void* host, uint64_t bytes) {
  uint64_t need = 0;
  int rc = hdsr_debug_size(c, which, &need);
  if (rc) return rc;
  if (!host && need) return HDSR_E_INVALID_ARG;
  if (bytes < need) return fail(c, HDSR_E_INVALID_ARG, "host buffer too small");
  if (need == 0) return HDSR_OK;
  if (which == HDSR_DBG_SORTED) {
    HIPX(c, hipDeviceSynchronize());
    HIPX(c, hipMemcpy(host, c->sorted, need, hipMemcpyDeviceToHost));
    return HDSR_OK;
  }
  const Bufs& b = c->eng.bufs();
  const DevStats& s = c->eng.host_stats();
  switch (which) {
    case HDSR_DBG_PROJCACHE: {
      uint32_t* tmp = nullptr;
      HIPX(c, hipMalloc((void**)&tmp, need));
      launch_cache14(b, 0, c->n, tmp, nullptr);
      const hipError_t e = hipMemcpy(host, tmp, need, hipMemcpyDeviceToHost);
      (void)hipFree(tmp);
      HIPX(c, e);
      return HDSR_OK;
    }
    case HDSR_DBG_DEPTH: HIPX(c, hipMemcpy(host, b.depth, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_TILE_COUNTS: HIPX(c, hipMemcpy(host, b.tile_counts, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_TILE_OFFSETS: HIPX(c, hipMemcpy(host, b.tile_offsets, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_ENTRIES:
      // a frame that sorted only the chunks its pixels reach gets the others sorted first
      if (int e = c->eng.complete_lists(nullptr)) return fail(c, e, c->eng.err);
      HIPX(c, hipMemcpy(host, b.entries, need, hipMemcpyDeviceToHost));
      return HDSR_OK;
    case HDSR_DBG_ENTRIES_CANON: HIPX(c, hipMemcpy(host, b.entries_canon, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_KEYS20: HIPX(c, hipMemcpy(host, b.keys20, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_SMALL_LIST: HIPX(c, hipMemcpy(host, b.small_list, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_LARGE_LIST: HIPX(c, hipMemcpy(host, b.large_list_t, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_LARGE_BASES: HIPX(c, hipMemcpy(host, b.large_bases, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_LARGE_SPLATS: HIPX(c, hipMemcpy(host, b.large_list, need, hipMemcpyDeviceToHost)); return HDSR_OK;
    case HDSR_DBG_CHUNKS: {
      std::vector<Seg> segs(s.segments);
      HIPX(c, hipMemcpy(segs.data(), b.segs, segs.size() * sizeof(Seg), hipMemcpyDeviceToHost));
      uint32_t* o = static_cast<uint32_t*>(host);
      size_t k = 0;
      for (const Seg& g : segs)
        if (g.large != 0xFFFFFFFFu) {
          o[k++] = g.dst;
          o[k++] = g.count & ~kSegKeepOrder;
        }
      return HDSR_OK;
    }
    default: return fail(c, HDSR_E_INVALID_ARG, "unknown debug buffer");
  }
}

}  // extern "C"
