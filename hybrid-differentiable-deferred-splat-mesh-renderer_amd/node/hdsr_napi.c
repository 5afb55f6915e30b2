/* hdsr_napi.c - Node N-API addon over the C ABI of include/hdsr.h (and the PLY decoder of
 * include/hdsr_assets.h).
 *
 * Thin: every function converts JS values (plain objects, typed arrays, an external handle) to
 * the C structs and calls exactly one hdsr_* entry point; a nonzero return becomes a thrown JS
 * Error carrying hdsr_last_error().  The handle's finalizer calls hdsr_destroy, so a compositor
 * that is garbage-collected releases its device memory.  Built with N-API version 8 headers
 * (Node >= 12.22).  The JavaScript facade with the reference's names is gpuTileSplatCompositor.js.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsr.h"
#include "hdsr_assets.h"

#define CALL(env, call)                                        \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                             \
    }                                                          \
  } while (0)

static napi_value throw_rc(napi_env env, int rc, hdsr_ctx* ctx, const char* what) {
  char buf[512];
  /* with no context, only hdsr_create leaves a message (hdsr_last_error(NULL)) */
  const char* msg = ctx ? hdsr_last_error(ctx) : strcmp(what, "hdsr_create") == 0 ? hdsr_last_error(NULL) : "";
  snprintf(buf, sizeof buf, "%s failed (%d): %s", what, rc, msg);
  napi_throw_error(env, NULL, buf);
  return NULL;
}

static uint32_t get_u32(napi_env env, napi_value obj, const char* key, uint32_t dflt) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  uint32_t out = dflt;
  napi_get_value_uint32(env, v, &out);
  return out;
}

static double get_f64(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  double out = dflt;
  napi_get_value_double(env, v, &out);
  return out;
}

static void set_u32(napi_env env, napi_value obj, const char* key, uint32_t v) {
  napi_value x;
  napi_create_uint32(env, v, &x);
  napi_set_named_property(env, obj, key, x);
}

static void set_f64(napi_env env, napi_value obj, const char* key, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, key, x);
}

/* Pointer and byte length of a typed array property (NULL when absent or undefined). */
static const void* typed_prop(napi_env env, napi_value obj, const char* key, size_t* bytes) {
  bool has = false;
  napi_value v;
  *bytes = 0;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return NULL;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return NULL;
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return NULL;
  napi_typedarray_type type;
  size_t length, offset;
  void* data;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &offset) != napi_ok) return NULL;
  size_t elem = 1;
  switch (type) {
    case napi_int16_array:
    case napi_uint16_array: elem = 2; break;
    case napi_int32_array:
    case napi_uint32_array:
    case napi_float32_array: elem = 4; break;
    case napi_float64_array:
    case napi_bigint64_array:
    case napi_biguint64_array: elem = 8; break;
    default: elem = 1;
  }
  *bytes = length * elem;
  return data;
}

static void plan_to_js(napi_env env, const hdsr_plan* p, napi_value out) {
  set_u32(env, out, "viewportWidth", p->viewport_width);
  set_u32(env, out, "viewportHeight", p->viewport_height);
  set_u32(env, out, "tileSizePx", p->tile_size_px);
  set_u32(env, out, "tileColumns", p->tile_columns);
  set_u32(env, out, "tileRows", p->tile_rows);
  set_u32(env, out, "tileCount", p->tile_count);
  set_u32(env, out, "mortonTileCount", p->morton_tile_count);
  set_u32(env, out, "splatCount", p->splat_count);
  set_u32(env, out, "maxTotalTileRefs", p->max_total_tile_refs);
}

static void plan_from_js(napi_env env, napi_value obj, hdsr_plan* p) {
  p->viewport_width = get_u32(env, obj, "viewportWidth", 0);
  p->viewport_height = get_u32(env, obj, "viewportHeight", 0);
  p->tile_size_px = get_u32(env, obj, "tileSizePx", 0);
  p->tile_columns = get_u32(env, obj, "tileColumns", 0);
  p->tile_rows = get_u32(env, obj, "tileRows", 0);
  p->tile_count = get_u32(env, obj, "tileCount", 0);
  p->morton_tile_count = get_u32(env, obj, "mortonTileCount", 0);
  p->splat_count = get_u32(env, obj, "splatCount", 0);
  p->max_total_tile_refs = get_u32(env, obj, "maxTotalTileRefs", 0);
}

/* The external holds a box so that destroy() can release the context early and the finalizer
 * (run at garbage collection) then only frees the box.  While a renderAsync() frame is in flight
 * the worker thread owns the context: 'busy' makes every other call on it throw, and a destroy()
 * in that window is deferred to the frame's completion ('destroy_pending'). */
typedef struct {
  hdsr_ctx* ctx;
  int busy;
  int destroy_pending;
} ctx_box;

static ctx_box* box_arg(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a compositor handle");
    return NULL;
  }
  return (ctx_box*)p;
}

static hdsr_ctx* ctx_arg(napi_env env, napi_value v) {
  ctx_box* b = box_arg(env, v);
  if (!b) return NULL;
  if (!b->ctx || b->destroy_pending) {
    napi_throw_error(env, NULL, "compositor handle already destroyed");
    return NULL;
  }
  if (b->busy) {
    napi_throw_error(env, NULL, "compositor is busy: a renderAsync() frame is in flight; wait for its Promise");
    return NULL;
  }
  return b->ctx;
}

static void finalize_ctx(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  ctx_box* b = (ctx_box*)data;
  if (b && b->ctx) hdsr_destroy(b->ctx);
  free(b);
}

/* planCompute({viewportWidth, viewportHeight, tileSizePx, splatCount, tileEntryMultiplier?}) */
static napi_value js_plan(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) {
    napi_throw_type_error(env, NULL, "planCompute(input)");
    return NULL;
  }
  hdsr_plan_in in;
  in.viewport_width = get_u32(env, argv[0], "viewportWidth", 0);
  in.viewport_height = get_u32(env, argv[0], "viewportHeight", 0);
  in.tile_size_px = get_u32(env, argv[0], "tileSizePx", 0);
  in.splat_count = get_u32(env, argv[0], "splatCount", 0);
  in.tile_entry_multiplier = (float)get_f64(env, argv[0], "tileEntryMultiplier", 0.0);
  hdsr_plan p;
  const int rc = hdsr_plan_compute(&in, &p);
  if (rc) return throw_rc(env, rc, NULL, "hdsr_plan_compute");
  napi_value out;
  CALL(env, napi_create_object(env, &out));
  plan_to_js(env, &p, out);
  return out;
}

/* create(plan, flags) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) {
    napi_throw_type_error(env, NULL, "create(plan, flags)");
    return NULL;
  }
  hdsr_plan p;
  plan_from_js(env, argv[0], &p);
  uint32_t flags = 0;
  if (argc > 1) napi_get_value_uint32(env, argv[1], &flags);
  hdsr_ctx* ctx = NULL;
  const int rc = hdsr_create(&p, flags, &ctx);
  if (rc) return throw_rc(env, rc, NULL, "hdsr_create");
  ctx_box* box = (ctx_box*)malloc(sizeof(ctx_box));
  if (!box) {
    hdsr_destroy(ctx);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  box->ctx = ctx;
  box->busy = 0;
  box->destroy_pending = 0;
  napi_value h;
  if (napi_create_external(env, box, finalize_ctx, NULL, &h) != napi_ok) {
    hdsr_destroy(ctx);
    free(box);
    napi_throw_error(env, NULL, "napi_create_external failed");
    return NULL;
  }
  return h;
}

/* destroy(handle): releases now, or - while a renderAsync() frame is in flight - as soon as that
 * frame completes; the handle must not be used again either way. */
static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ctx_box* b = box_arg(env, argv[0]);
  if (!b) return NULL;
  if (b->busy) {
    b->destroy_pending = 1;
    return NULL;
  }
  if (b->ctx) hdsr_destroy(b->ctx);
  b->ctx = NULL;
  return NULL;
}

/* Element count of typed array property 'key' against 'want' elements ('exact': equal, else at
 * least); throws a RangeError naming the property on a mismatch.  Returns 0 when it threw. */
static int check_len(napi_env env, const char* key, size_t bytes, size_t elem, uint64_t want, int exact) {
  const uint64_t have = bytes / elem;
  if (exact ? have == want : have >= want) return 1;
  char buf[256];
  snprintf(buf, sizeof buf, "%s holds %llu elements; %s%llu expected for the plan's splatCount", key,
           (unsigned long long)have, exact ? "" : "at least ", (unsigned long long)want);
  napi_throw_range_error(env, NULL, buf);
  return 0;
}

/* bindSplats(handle, count, {positions, scales, rotations, opacities, materials, normals?, shData,
 *                            sortedIndices, shDegree?})  -- host typed arrays, copied to the
 * device.  Every array's length is checked against 'count' (the plan's splatCount) before
 * anything is read: positions / scales 3N, rotations 4N, opacities / materials / sortedIndices N,
 * shData N * (6 + 3K); normals any length (reads are bounded by it).  Without shDegree the degree is taken from the
 * shData length, as the reference's bind names no degree (src/splatRenderer.ts:946). */
static napi_value js_bind(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3) {
    napi_throw_type_error(env, NULL, "bindSplats(handle, count, buffers)");
    return NULL;
  }
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  uint32_t n = 0;
  if (napi_get_value_uint32(env, argv[1], &n) != napi_ok) {
    napi_throw_type_error(env, NULL, "bindSplats: count must be a number");
    return NULL;
  }
  size_t b = 0;
  hdsr_splats s;
  memset(&s, 0, sizeof s);
  s.count = n;
  s.positions = (const float*)typed_prop(env, argv[2], "positions", &b);
  if (!check_len(env, "positionBuffer", b, 4, 3ull * n, 1)) return NULL;
  s.scales = (const float*)typed_prop(env, argv[2], "scales", &b);
  if (!check_len(env, "scaleBuffer", b, 4, 3ull * n, 1)) return NULL;
  s.rotations = (const float*)typed_prop(env, argv[2], "rotations", &b);
  if (!check_len(env, "rotationBuffer", b, 4, 4ull * n, 1)) return NULL;
  s.opacities = (const float*)typed_prop(env, argv[2], "opacities", &b);
  if (!check_len(env, "opacityBuffer", b, 4, n, 1)) return NULL;
  s.materials = (const uint32_t*)typed_prop(env, argv[2], "materials", &b);
  if (!check_len(env, "materialBuffer", b, 4, n, 1)) return NULL;
  s.sorted_indices = (const uint32_t*)typed_prop(env, argv[2], "sortedIndices", &b);
  if (!check_len(env, "sortedIndexBuffer", b, 4, n, 1)) return NULL;
  /* any length: the projection reads a splat's normal only where the buffer holds it, as the
   * shader bounds its reads by arrayLength (gpu_project_splats.wgsl:237,252) */
  s.normals = (const float*)typed_prop(env, argv[2], "normals", &b);
  s.normals_len = b / 4;
  s.sh_data = (const float*)typed_prop(env, argv[2], "shData", &b);
  s.sh_data_len = b / 4;
  s.sh_degree = get_u32(env, argv[2], "shDegree", HDSR_SH_DEGREE_AUTO);
  if (s.sh_degree == HDSR_SH_DEGREE_AUTO && n) {
    int found = 0;
    for (uint64_t d = 0; d <= 3; ++d) found |= s.sh_data_len == (6 + 3 * ((d + 1) * (d + 1) - 1)) * (uint64_t)n;
    if (!found) {
      char buf[256];
      snprintf(buf, sizeof buf,
               "shDataBuffer holds %llu floats, which matches no SH degree: %u splats need 6, 15, 30 or 51 each",
               (unsigned long long)s.sh_data_len, n);
      napi_throw_range_error(env, NULL, buf);
      return NULL;
    }
  }
  if (s.sh_degree != HDSR_SH_DEGREE_AUTO && s.sh_degree <= 3) {
    const uint64_t k = (uint64_t)(s.sh_degree + 1) * (s.sh_degree + 1) - 1;
    if (!check_len(env, "shDataBuffer", b, 4, (6 + 3 * k) * (uint64_t)n, 0)) return NULL;
  }
  s.memory = HDSR_MEM_HOST;
  const int rc = hdsr_bind_splats(ctx, &s);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_bind_splats");
  return NULL;
}

/* bindOutputs(handle, rgba, depth): caller-owned device buffers (BigInt device addresses, or
 * null for the context's own) that later renders write - hdsr_bind_outputs. */
static napi_value js_bind_outputs(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  void* ptrs[2] = {NULL, NULL};
  for (int i = 0; i < 2; ++i) {
    if (argc <= (size_t)(i + 1)) break;
    napi_valuetype t;
    napi_typeof(env, argv[i + 1], &t);
    if (t == napi_null || t == napi_undefined) continue;
    if (t != napi_bigint) {
      napi_throw_type_error(env, NULL, "bindOutputs: device addresses must be BigInt (or null)");
      return NULL;
    }
    uint64_t v = 0;
    bool lossless = false;
    CALL(env, napi_get_value_bigint_uint64(env, argv[i + 1], &v, &lossless));
    ptrs[i] = (void*)(uintptr_t)v;
  }
  const int rc = hdsr_bind_outputs(ctx, (float*)ptrs[0], (float*)ptrs[1]);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_bind_outputs");
  return NULL;
}

/* outputs(handle) -> {rgba, depth, gbufferNormal, gbufferMaterial, rgba16} device addresses
 * (BigInt, 0n when absent), width, height - hdsr_get_outputs. */
static napi_value js_outputs(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  hdsr_outputs o;
  const int rc = hdsr_get_outputs(ctx, &o);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_get_outputs");
  napi_value out, v;
  CALL(env, napi_create_object(env, &out));
  const struct {
    const char* key;
    const void* p;
  } f[] = {{"rgba", o.rgba}, {"depth", o.depth}, {"gbufferNormal", o.gbuffer_normal},
           {"gbufferMaterial", o.gbuffer_material}, {"rgba16", o.rgba16}, {"lit", o.lit}};
  for (size_t i = 0; i < sizeof f / sizeof f[0]; ++i) {
    CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)f[i].p, &v));
    CALL(env, napi_set_named_property(env, out, f[i].key, v));
  }
  set_u32(env, out, "width", o.width);
  set_u32(env, out, "height", o.height);
  return out;
}

/* updateSortedIndices(handle, Uint32Array, count): hdsr_update_sorted_indices after a length
 * check against count (the plan's splatCount). */
static napi_value js_update_sorted(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  napi_value holder;
  CALL(env, napi_create_object(env, &holder));
  CALL(env, napi_set_named_property(env, holder, "a", argv[1]));
  size_t bytes = 0;
  const void* p = typed_prop(env, holder, "a", &bytes);
  uint32_t n = 0;
  if (argc < 3 || napi_get_value_uint32(env, argv[2], &n) != napi_ok) {
    napi_throw_type_error(env, NULL, "updateSortedIndices(handle, sortedIndices, count)");
    return NULL;
  }
  if (!p || !check_len(env, "sortedIndices", bytes, 4, n, 1)) {
    if (!p) napi_throw_type_error(env, NULL, "sortedIndices must be a Uint32Array");
    return NULL;
  }
  const int rc = hdsr_update_sorted_indices(ctx, (const uint32_t*)p, HDSR_MEM_HOST);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_update_sorted_indices");
  return NULL;
}

/* setFrame(handle, Float32Array(64) | ArrayBuffer view of 256 bytes) */
static napi_value js_set_frame(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  napi_value holder;
  CALL(env, napi_create_object(env, &holder));
  CALL(env, napi_set_named_property(env, holder, "a", argv[1]));
  size_t bytes = 0;
  const void* p = typed_prop(env, holder, "a", &bytes);
  if (!p || bytes < HDSR_FRAME_UNIFORM_BYTES) {
    napi_throw_type_error(env, NULL, "frame uniforms must be a 256-byte typed array");
    return NULL;
  }
  const int rc = hdsr_set_frame(ctx, p);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_set_frame");
  return NULL;
}

static napi_value js_render(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  const int rc = hdsr_render(ctx, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_render");
  return NULL;
}

/* renderAsync(h): hdsr_render and hdsr_sync on a libuv worker thread; returns a Promise that
 * resolves once the frame is complete (capacity re-runs included) or rejects with the error.
 * The context must not be used from JS until the Promise settles: a context serves one host
 * thread at a time. */
typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref handle_ref; /* keeps the context's external alive while the worker uses it */
  ctx_box* box;
  hdsr_ctx* ctx;
  int rc;
  const char* what;
  char err[512];
} render_job;

static void render_execute(napi_env env, void* data) {
  render_job* j = (render_job*)data;
  (void)env;
  j->what = "hdsr_render";
  j->rc = hdsr_render(j->ctx, NULL);
  if (!j->rc) {
    j->what = "hdsr_sync";
    j->rc = hdsr_sync(j->ctx);
  }
  if (j->rc) snprintf(j->err, sizeof j->err, "%s", hdsr_last_error(j->ctx));
}

static void render_complete(napi_env env, napi_status status, void* data) {
  render_job* j = (render_job*)data;
  j->box->busy = 0;
  if (j->box->destroy_pending) {  /* destroy() was called while the frame ran */
    if (j->box->ctx) hdsr_destroy(j->box->ctx);
    j->box->ctx = NULL;
  }
  if (status == napi_ok && j->rc == 0) {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, j->deferred, undef);
  } else {
    char buf[640];
    napi_value msg, err;
    if (status != napi_ok)
      snprintf(buf, sizeof buf, "renderAsync: the work was not completed (status %d)", (int)status);
    else
      snprintf(buf, sizeof buf, "%s failed (%d): %s", j->what, j->rc, j->err);
    napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_reference(env, j->handle_ref);
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value js_render_async(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], promise, name;
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  render_job* j = (render_job*)calloc(1, sizeof *j);
  if (!j) {
    napi_throw_error(env, NULL, "renderAsync: out of host memory");
    return NULL;
  }
  j->ctx = ctx;
  j->box = box_arg(env, argv[0]);
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_reference(env, argv[0], 1, &j->handle_ref) != napi_ok ||
      napi_create_string_utf8(env, "hdsr.renderAsync", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, render_execute, render_complete, j, &j->work) != napi_ok) {
    free(j);
    napi_throw_error(env, NULL, "renderAsync: N-API setup failed");
    return NULL;
  }
  j->box->busy = 1;
  if (napi_queue_async_work(env, j->work) != napi_ok) {
    j->box->busy = 0;
    napi_delete_reference(env, j->handle_ref);
    napi_delete_async_work(env, j->work);
    free(j);
    napi_throw_error(env, NULL, "renderAsync: could not queue the work");
    return NULL;
  }
  return promise;
}

static napi_value js_sort_back_to_front(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  const int rc = hdsr_sort_back_to_front(ctx, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_sort_back_to_front");
  return NULL;
}

static napi_value js_composite_only(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  const int rc = hdsr_composite_only(ctx, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_composite_only");
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  const int rc = hdsr_sync(ctx);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_sync");
  return NULL;
}

/* readback(handle, which, typedArray) copies an output into the array */
static napi_value js_readback(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  int32_t which = 0;
  napi_get_value_int32(env, argv[1], &which);
  napi_value holder;
  CALL(env, napi_create_object(env, &holder));
  CALL(env, napi_set_named_property(env, holder, "a", argv[2]));
  size_t bytes = 0;
  void* p = (void*)typed_prop(env, holder, "a", &bytes);
  if (!p) {
    napi_throw_type_error(env, NULL, "readback: the destination must be a typed array");
    return NULL;
  }
  const int rc = hdsr_readback(ctx, which, p, bytes);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_readback");
  return NULL;
}

/* debugReadback(handle, which) -> Uint32Array */
static napi_value js_debug(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  int32_t which = 0;
  napi_get_value_int32(env, argv[1], &which);
  uint64_t bytes = 0;
  int rc = hdsr_debug_size(ctx, which, &bytes);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_debug_size");
  void* data = NULL;
  napi_value ab, ta;
  CALL(env, napi_create_arraybuffer(env, (size_t)bytes, &data, &ab));
  rc = hdsr_debug_readback(ctx, which, data, bytes);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_debug_readback");
  CALL(env, napi_create_typedarray(env, napi_uint32_array, (size_t)(bytes / 4), ab, 0, &ta));
  return ta;
}

static napi_value js_stats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  hdsr_stats s;
  const int rc = hdsr_get_stats(ctx, &s);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_get_stats");
  napi_value o;
  CALL(env, napi_create_object(env, &o));
  set_u32(env, o, "refs", s.refs);
  set_u32(env, o, "smallPairs", s.small_pairs);
  set_u32(env, o, "largePairs", s.large_pairs);
  set_u32(env, o, "largeSplats", s.large_splats);
  set_u32(env, o, "visibleSplats", s.visible_splats);
  set_u32(env, o, "smallTiles", s.small_tiles);
  set_u32(env, o, "largeTiles", s.large_tiles);
  set_u32(env, o, "chunks", s.chunks);
  set_u32(env, o, "segments", s.segments);
  set_u32(env, o, "pairCapacity", s.pair_capacity);
  set_u32(env, o, "overflowReruns", s.overflow_reruns);
  set_u32(env, o, "capRefsTripped", s.cap_refs_tripped);
  set_u32(env, o, "capLargeSplatsTripped", s.cap_large_splats_tripped);
  set_u32(env, o, "capChunksTripped", s.cap_chunks_tripped);
  set_u32(env, o, "compositeRefs", s.composite_refs);
  set_u32(env, o, "compositeMaxRefs", s.composite_max_refs);
  set_u32(env, o, "largeTileRefs", s.large_tile_refs);
  set_u32(env, o, "cellPairs", s.cell_pairs);
  set_u32(env, o, "sortedRefs", s.sorted_refs);
  set_u32(env, o, "resumedTiles", s.resumed_tiles);
  set_u32(env, o, "resumedTileRefs", s.resumed_tile_refs);
  set_u32(env, o, "firstChunks", s.first_chunks);
  return o;
}

static napi_value js_stage_times(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  float ms[HDSR_STAGE_COUNT];
  uint32_t n = 0;
  const int rc = hdsr_stage_times(ctx, ms, &n);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_stage_times");
  static const char* names[HDSR_STAGE_COUNT] = {"project", "scan", "emit", "group", "ranges",
                                                "classify", "bucket", "sort", "composite"};
  napi_value o;
  CALL(env, napi_create_object(env, &o));
  for (int i = 0; i < HDSR_STAGE_COUNT; ++i) set_f64(env, o, names[i], ms[i]);
  set_u32(env, o, "renders", n);
  return o;
}

/* New Float32Array / Uint32Array of n elements; *data receives its storage. */
static napi_value new_typed(napi_env env, napi_typedarray_type type, size_t n, void** data) {
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * 4, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, type, n, ab, 0, &ta) != napi_ok) return NULL;
  if (n) memset(*data, 0, n * 4);
  return ta;
}

static napi_value vec3(napi_env env, const double* v) {
  napi_value a, x;
  napi_create_array_with_length(env, 3, &a);
  for (uint32_t i = 0; i < 3; ++i) {
    napi_create_double(env, v[i], &x);
    napi_set_element(env, a, i, x);
  }
  return a;
}

/* plyDecode(fileName, bytes: ArrayBuffer | typed array) -> { count, shDegree, positions, ...,
 * bounds }: hdsr_ply_inspect + hdsr_ply_decode; a failure throws the decoder's message. */
static napi_value js_ply_decode(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  char name[1024] = {0};
  size_t nl = 0;
  napi_get_value_string_utf8(env, argv[0], name, sizeof name, &nl);
  const void* bytes = NULL;
  size_t len = 0;
  bool is_ab = false, is_ta = false;
  napi_is_arraybuffer(env, argv[1], &is_ab);
  napi_is_typedarray(env, argv[1], &is_ta);
  if (is_ab) {
    void* d;
    CALL(env, napi_get_arraybuffer_info(env, argv[1], &d, &len));
    bytes = d;
  } else if (is_ta) {
    napi_value holder;
    CALL(env, napi_create_object(env, &holder));
    CALL(env, napi_set_named_property(env, holder, "a", argv[1]));
    bytes = typed_prop(env, holder, "a", &len);
  }
  if (!bytes) {
    napi_throw_type_error(env, NULL, "plyDecode: bytes must be an ArrayBuffer or a typed array");
    return NULL;
  }
  char err[1024] = {0};
  hdsr_ply_info pi;
  int rc = hdsr_ply_inspect(name, bytes, len, &pi, err, sizeof err);
  if (rc) {
    napi_throw_error(env, NULL, err);
    return NULL;
  }
  const size_t n = pi.count, k = pi.sh_coefficient_count;
  hdsr_splat_attributes o;
  memset(&o, 0, sizeof o);
  napi_value res;
  CALL(env, napi_create_object(env, &res));
  struct {
    const char* key;
    void** dst;
    size_t len;
    napi_typedarray_type type;
    int present;
  } outs[] = {
      {"positions", (void**)&o.positions, n * 3, napi_float32_array, 1},
      {"colors", (void**)&o.colors, n * 3, napi_float32_array, 1},
      {"opacities", (void**)&o.opacities, n, napi_float32_array, 1},
      {"radii", (void**)&o.radii, n, napi_float32_array, 1},
      {"scales", (void**)&o.scales, n * 3, napi_float32_array, 1},
      {"rotations", (void**)&o.rotations, n * 4, napi_float32_array, 1},
      {"shCoefficients", (void**)&o.sh_coefficients, n * k * 3, napi_float32_array, k != 0},
      {"normals", (void**)&o.normals, n * 3, napi_float32_array, (pi.fields & HDSR_PLY_NORMALS) != 0},
      {"roughness", (void**)&o.roughness, n, napi_float32_array, (pi.fields & HDSR_PLY_ROUGHNESS) != 0},
      {"metalness", (void**)&o.metalness, n, napi_float32_array, (pi.fields & HDSR_PLY_METALNESS) != 0},
      {"emissive", (void**)&o.emissive, n * 3, napi_float32_array, (pi.fields & HDSR_PLY_EMISSIVE) != 0},
      {"detailNormals", (void**)&o.detail_normals, n * 3, napi_float32_array,
       (pi.fields & HDSR_PLY_DETAIL_NORMALS) != 0},
      {"originalIds", (void**)&o.original_ids, n, napi_uint32_array, 1},
  };
  for (size_t i = 0; i < sizeof outs / sizeof outs[0]; ++i) {
    if (!outs[i].present) continue;
    napi_value ta = new_typed(env, outs[i].type, outs[i].len, outs[i].dst);
    if (!ta) {
      napi_throw_error(env, NULL, "plyDecode: allocation failed");
      return NULL;
    }
    CALL(env, napi_set_named_property(env, res, outs[i].key, ta));
  }
  rc = hdsr_ply_decode(name, bytes, len, &o, err, sizeof err);
  if (rc) {
    napi_throw_error(env, NULL, err);
    return NULL;
  }
  set_u32(env, res, "count", pi.count);
  set_u32(env, res, "shDegree", pi.sh_degree);
  set_u32(env, res, "shCoefficientCount", pi.sh_coefficient_count);
  napi_value bounds;
  CALL(env, napi_create_object(env, &bounds));
  CALL(env, napi_set_named_property(env, bounds, "min", vec3(env, o.bounds_min)));
  CALL(env, napi_set_named_property(env, bounds, "max", vec3(env, o.bounds_max)));
  CALL(env, napi_set_named_property(env, bounds, "center", vec3(env, o.bounds_center)));
  set_f64(env, bounds, "radius", o.bounds_radius);
  CALL(env, napi_set_named_property(env, res, "bounds", bounds));
  return res;
}

/* The reference's deferred-normal parameters, overridden by the fields a JS object sets. */
static void normal_params_from(napi_env env, napi_value obj, hdsr_normal_params* p) {
  hdsr_normal_params_default(p);
  p->recovery_mode = get_u32(env, obj, "recoveryMode", p->recovery_mode);
  p->grazing_start = (float)get_f64(env, obj, "grazingStart", p->grazing_start);
  p->grazing_end = (float)get_f64(env, obj, "grazingEnd", p->grazing_end);
  p->spatial_sigma = (float)get_f64(env, obj, "spatialSigma", p->spatial_sigma);
  p->depth_sigma = (float)get_f64(env, obj, "depthSigma", p->depth_sigma);
  p->normal_sigma = (float)get_f64(env, obj, "normalSigma", p->normal_sigma);
  p->kernel_radius = get_u32(env, obj, "kernelRadius", p->kernel_radius);
}

static napi_value js_screen_space_normals(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  hdsr_normal_params p;
  normal_params_from(env, argv[1], &p);
  const int rc = hdsr_screen_space_normals(ctx, &p, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_screen_space_normals");
  return NULL;
}

/* A property holding a number array (plain or typed) of exactly n elements, into out; false when
 * absent.  A present property of another shape throws a TypeError (returns -1). */
static int get_vec(napi_env env, napi_value obj, const char* key, double* out, uint32_t n) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return 0;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_undefined || t == napi_null) return 0;
  bool is_arr = false, is_ta = false;
  napi_is_array(env, v, &is_arr);
  napi_is_typedarray(env, v, &is_ta);
  uint32_t len = 0;
  if (is_arr) {
    napi_get_array_length(env, v, &len);
  } else if (is_ta) {
    napi_typedarray_type tt;
    size_t l, off;
    void* d;
    napi_value ab;
    napi_get_typedarray_info(env, v, &tt, &l, &d, &ab, &off);
    len = (uint32_t)l;
  }
  if ((!is_arr && !is_ta) || len != n) {
    char msg[128];
    snprintf(msg, sizeof msg, "%s must be an array of %u numbers", key, n);
    napi_throw_type_error(env, NULL, msg);
    return -1;
  }
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    napi_get_element(env, v, i, &e);
    if (napi_get_value_double(env, e, &out[i]) != napi_ok) {
      char msg[128];
      snprintf(msg, sizeof msg, "%s must be an array of %u numbers", key, n);
      napi_throw_type_error(env, NULL, msg);
      return -1;
    }
  }
  return 1;
}

static void curve_from(napi_env env, napi_value obj, const char* key, hdsr_material_curve* c) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_object) return;
  c->contrast = get_f64(env, v, "contrast", c->contrast);
  c->brightness = get_f64(env, v, "brightness", c->brightness);
  c->gamma = get_f64(env, v, "gamma", c->gamma);
}

/* A BigInt device address property (0 when absent / null); -1 for another type (TypeError). */
static int get_devptr(napi_env env, napi_value obj, const char* key, const float** out) {
  bool has = false;
  napi_value v;
  *out = NULL;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return 0;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_undefined || t == napi_null) return 0;
  if (t != napi_bigint) {
    char msg[128];
    snprintf(msg, sizeof msg, "%s must be a device address (BigInt) or null", key);
    napi_throw_type_error(env, NULL, msg);
    return -1;
  }
  uint64_t a = 0;
  bool lossless = false;
  napi_get_value_bigint_uint64(env, v, &a, &lossless);
  *out = (const float*)(uintptr_t)a;
  return 1;
}

/* deferredLighting(handle, options): the encode arguments of createDeferredLightingPass
 * (src/splatRenderer.ts:484-506) by name; unset fields take the reference viewer's values. */
static napi_value js_deferred_lighting(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  hdsr_lighting_params p;
  hdsr_lighting_params_default(&p);
  napi_valuetype t = napi_undefined;
  if (argc > 1) napi_typeof(env, argv[1], &t);
  float vp[16], cam[3];
  if (t == napi_object) {
    const napi_value o = argv[1];
    if (get_vec(env, o, "lightDirection", p.light_direction, 3) < 0) return NULL;
    p.light_intensity = get_f64(env, o, "lightIntensity", p.light_intensity);
    p.ambient_intensity = get_f64(env, o, "ambientIntensity", p.ambient_intensity);
    bool has = false;
    napi_value sv;
    if (napi_has_named_property(env, o, "specularOnly", &has) == napi_ok && has &&
        napi_get_named_property(env, o, "specularOnly", &sv) == napi_ok) {
      bool b = false;
      if (napi_get_value_bool(env, sv, &b) == napi_ok) p.specular_only = b ? 1u : 0u;
    }
    p.emissive_intensity = get_f64(env, o, "emissiveIntensity", p.emissive_intensity);
    p.emissive_threshold = get_f64(env, o, "emissiveThreshold", p.emissive_threshold);
    p.env_intensity = get_f64(env, o, "envIntensity", p.env_intensity);
    p.env_rotation = get_f64(env, o, "envRotation", p.env_rotation);
    p.exposure = get_f64(env, o, "exposure", p.exposure);
    curve_from(env, o, "roughnessCurve", &p.roughness_curve);
    curve_from(env, o, "metalnessCurve", &p.metalness_curve);
    curve_from(env, o, "albedoCurve", &p.albedo_curve);
    double d16[16], d3[3];
    int r = get_vec(env, o, "lightingViewProj", d16, 16);
    if (r < 0) return NULL;
    if (r > 0) {
      for (int i = 0; i < 16; ++i) vp[i] = (float)d16[i];
      p.lighting_view_proj = vp;
    }
    r = get_vec(env, o, "lightingCameraPosition", d3, 3);
    if (r < 0) return NULL;
    if (r > 0) {
      for (int i = 0; i < 3; ++i) cam[i] = (float)d3[i];
      p.lighting_camera_position = cam;
    }
    if (get_devptr(env, o, "ao", &p.ao) < 0 || get_devptr(env, o, "bloom", &p.bloom) < 0) return NULL;
  }
  const int rc = hdsr_deferred_lighting(ctx, &p, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_deferred_lighting");
  return NULL;
}

static napi_value js_bilateral_normal_filter(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  hdsr_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return NULL;
  hdsr_normal_params p;
  normal_params_from(env, argv[1], &p);
  const int rc = hdsr_bilateral_normal_filter(ctx, &p, NULL);
  if (rc) return throw_rc(env, rc, ctx, "hdsr_bilateral_normal_filter");
  return NULL;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value s;
  CALL(env, napi_create_string_utf8(env, hdsr_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

static napi_value init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"planCompute", NULL, js_plan, NULL, NULL, NULL, napi_default, NULL},
      {"create", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
      {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
      {"bindSplats", NULL, js_bind, NULL, NULL, NULL, napi_default, NULL},
      {"bindOutputs", NULL, js_bind_outputs, NULL, NULL, NULL, napi_default, NULL},
      {"outputs", NULL, js_outputs, NULL, NULL, NULL, napi_default, NULL},
      {"updateSortedIndices", NULL, js_update_sorted, NULL, NULL, NULL, napi_default, NULL},
      {"setFrame", NULL, js_set_frame, NULL, NULL, NULL, napi_default, NULL},
      {"render", NULL, js_render, NULL, NULL, NULL, napi_default, NULL},
      {"renderAsync", NULL, js_render_async, NULL, NULL, NULL, napi_default, NULL},
      {"compositeOnly", NULL, js_composite_only, NULL, NULL, NULL, napi_default, NULL},
      {"sortBackToFront", NULL, js_sort_back_to_front, NULL, NULL, NULL, napi_default, NULL},
      {"sync", NULL, js_sync, NULL, NULL, NULL, napi_default, NULL},
      {"readback", NULL, js_readback, NULL, NULL, NULL, napi_default, NULL},
      {"debugReadback", NULL, js_debug, NULL, NULL, NULL, napi_default, NULL},
      {"stats", NULL, js_stats, NULL, NULL, NULL, napi_default, NULL},
      {"stageTimes", NULL, js_stage_times, NULL, NULL, NULL, napi_default, NULL},
      {"version", NULL, js_version, NULL, NULL, NULL, napi_default, NULL},
      {"plyDecode", NULL, js_ply_decode, NULL, NULL, NULL, napi_default, NULL},
      {"screenSpaceNormals", NULL, js_screen_space_normals, NULL, NULL, NULL, napi_default, NULL},
      {"bilateralNormalFilter", NULL, js_bilateral_normal_filter, NULL, NULL, NULL, napi_default, NULL},
      {"deferredLighting", NULL, js_deferred_lighting, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
