/* hdsr_assets.h - on-disk splat asset decoding (host code in libhdsr.so, no GPU work).
 *
 * Replaces decodeLocalPlySplatPayload (src/localPly.ts:272-420; header parse :488-598,
 * property reads :640-686, framing bounds :688-721): a binary little-endian 3DGS / Scaniverse
 * PLY vertex block becomes the SplatAttributes arrays of src/splats.ts:38-60, which the
 * compositor's packers (uploadSplatAttributeBuffers, src/splats.ts:376-500) turn into the
 * buffers hdsr_bind_splats takes.  The first-smoke manifest + sidecar format
 * (decodeFirstSmokeSplatManifest, src/splats.ts:230-344) is JSON and is decoded by the host
 * facades (hdsr/assets.py, node/splatAssets.js) over plain little-endian float32 sidecars.
 *
 * Values follow the reference's arithmetic: every property is read as a double (JavaScript
 * number), derived values are computed in double and stored to float32 once:
 *   colors    = clamp01(0.5 + 0.28209479177387814 * f_dc_c), else clamp01(byte / 255), else 0.5
 *   opacities = 1 / (1 + exp(-opacity)), or 1 without an opacity property
 *   radii     = max(exp(scale_0), exp(scale_1), exp(scale_2)) of the stored float32 scales
 *   sh        = f_rest_* (channel-major in the file) reordered coefficient-major, rgb-interleaved
 * Framing bounds: 5th / 95th percentiles of up to 50,000 evenly strided positions when
 * count > 100, else the exact extents.
 *
 * Errors: a negative HDSR_E_* code (hdsr.h) and the reference's message in 'err' (truncated to
 * err_len bytes, NUL-terminated; 'err' may be null). */
#ifndef HDSR_ASSETS_H
#define HDSR_ASSETS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* optional attribute groups present in a PLY (hdsr_ply_info.fields) */
#define HDSR_PLY_NORMALS 1u         /* nx ny nz */
#define HDSR_PLY_ROUGHNESS 2u       /* roughness */
#define HDSR_PLY_METALNESS 4u       /* metallic */
#define HDSR_PLY_EMISSIVE 8u        /* emissive_r emissive_g emissive_b */
#define HDSR_PLY_DETAIL_NORMALS 16u /* detail_normal_x _y _z */

typedef struct hdsr_ply_info {
  uint32_t count;                 /* vertex rows */
  uint32_t sh_degree;             /* 0 when the file has no f_rest_* properties */
  uint32_t sh_coefficient_count;  /* (degree + 1)^2 - 1 */
  uint32_t fields;                /* HDSR_PLY_* */
  uint64_t data_offset;           /* first byte of the vertex block */
  uint32_t row_stride;            /* bytes per vertex row */
  uint32_t property_count;
} hdsr_ply_info;

/* Parses and validates the header (parseBinaryLittleEndianPlyHeader, src/localPly.ts:488-547, and
 * the checks of decodeLocalPlySplatPayload before its row loop).  'file_name' only appears in
 * messages. */
int hdsr_ply_inspect(const char* file_name, const void* bytes, uint64_t len, hdsr_ply_info* info, char* err,
                     uint32_t err_len);

/* Caller-allocated outputs, sized from hdsr_ply_info.  Optional groups the file lacks are not
 * written (their pointers may be null); a pointer for an absent group is left untouched. */
typedef struct hdsr_splat_attributes {
  float* positions;        /* count * 3 */
  float* colors;           /* count * 3, linear display rgb */
  float* opacities;        /* count */
  float* radii;            /* count */
  float* scales;           /* count * 3, log-space as stored */
  float* rotations;        /* count * 4, wxyz as stored (rot_0..rot_3) */
  float* sh_coefficients;  /* count * coefficient_count * 3, splat_coeff_rgb order */
  float* normals;          /* count * 3 */
  float* roughness;        /* count */
  float* metalness;        /* count */
  float* emissive;         /* count * 3 */
  float* detail_normals;   /* count * 3 */
  uint32_t* original_ids;  /* count: the row index */
  double bounds_min[3];    /* outputs: SplatBounds (src/splats.ts:24-29) */
  double bounds_max[3];
  double bounds_center[3];
  double bounds_radius;
} hdsr_splat_attributes;

/* Decodes every vertex row (decodeLocalPlySplatPayload, src/localPly.ts:272-420). */
int hdsr_ply_decode(const char* file_name, const void* bytes, uint64_t len, hdsr_splat_attributes* out, char* err,
                    uint32_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* HDSR_ASSETS_H */
