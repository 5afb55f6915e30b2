// assets.cpp - binary little-endian 3DGS / Scaniverse PLY -> splat attribute arrays (host code).
//
// Reference: decodeLocalPlySplatPayload (src/localPly.ts:272-420) with its header parser
// (:488-598), SH layout discovery (:429-486), property reads (:640-686) and framing bounds
// (:688-721).  The decoded arrays are the SplatAttributes of src/splats.ts:38-60.
//
// The reference runs in JavaScript: every property is read into a double, derived values are
// computed in double and rounded to float32 when stored.  The same is done here (no contraction;
// min / max propagate NaN as Math.min / Math.max do).  exp is the C library's; JavaScript engines
// ship their own, so an opacity or radius may differ from the browser's in the last float32 bit
// when the double results straddle a float32 rounding boundary.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "hdsr.h"
#include "hdsr_assets.h"

namespace {

enum class PType { Char, UChar, Short, UShort, Int, UInt, Float, Double };

struct Prop {
  std::string name;
  PType type;
  uint32_t offset;
};

struct Header {
  std::string format;
  double vertex_count = NAN;
  std::vector<Prop> props;
  uint32_t stride = 0;
  uint64_t data_offset = 0;
};

int fail(char* err, uint32_t err_len, int code, const std::string& msg) {
  if (err && err_len) {
    const size_t n = std::min<size_t>(msg.size(), err_len - 1);
    memcpy(err, msg.data(), n);
    err[n] = 0;
  }
  return code;
}

// Math.min / Math.max: NaN if either operand is NaN
double js_min(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (a < b ? a : b); }
double js_max(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (a > b ? a : b); }
double clamp01(double v) { return js_max(0.0, js_min(1.0, v)); }

// JavaScript's Number(token) for the tokens a PLY header holds: decimal or hexadecimal,
// the whole token or NaN.
double js_number(const std::string& s) {
  if (s.empty()) return 0.0;
  const char* b = s.c_str();
  char* e = nullptr;
  const double v = std::strtod(b, &e);
  if (e != b + s.size()) return NAN;
  return v;
}

std::string js_int_string(double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.0f", v);
  return buf;
}

bool normalize_type(const std::string& t, PType* out) {
  static const std::pair<const char*, PType> names[] = {
      {"char", PType::Char},     {"int8", PType::Char},    {"uchar", PType::UChar},    {"uint8", PType::UChar},
      {"short", PType::Short},   {"int16", PType::Short},  {"ushort", PType::UShort},  {"uint16", PType::UShort},
      {"int", PType::Int},       {"int32", PType::Int},    {"uint", PType::UInt},      {"uint32", PType::UInt},
      {"float", PType::Float},   {"float32", PType::Float}, {"double", PType::Double}, {"float64", PType::Double}};
  for (const auto& n : names)
    if (t == n.first) {
      *out = n.second;
      return true;
    }
  return false;
}

uint32_t type_size(PType t) {
  switch (t) {
    case PType::Char:
    case PType::UChar: return 1;
    case PType::Short:
    case PType::UShort: return 2;
    case PType::Double: return 8;
    default: return 4;
  }
}

double read_prop(const uint8_t* row, const Prop& p) {
  const uint8_t* q = row + p.offset;
  switch (p.type) {
    case PType::Char: return (double)(int8_t)q[0];
    case PType::UChar: return (double)q[0];
    case PType::Short: { int16_t v; memcpy(&v, q, 2); return (double)v; }
    case PType::UShort: { uint16_t v; memcpy(&v, q, 2); return (double)v; }
    case PType::Int: { int32_t v; memcpy(&v, q, 4); return (double)v; }
    case PType::UInt: { uint32_t v; memcpy(&v, q, 4); return (double)v; }
    case PType::Float: { float v; memcpy(&v, q, 4); return (double)v; }
    case PType::Double: { double v; memcpy(&v, q, 8); return v; }
  }
  return NAN;
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

std::vector<std::string> split_ws(const std::string& line) {
  std::vector<std::string> parts;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && is_space(line[i])) ++i;
    const size_t j = i;
    while (i < line.size() && !is_space(line[i])) ++i;
    if (i > j) parts.push_back(line.substr(j, i - j));
  }
  if (parts.empty()) parts.push_back("");
  return parts;
}

// parseBinaryLittleEndianPlyHeader (src/localPly.ts:488-547) and findHeaderEnd (:572-598)
int parse_header(const char* name, const uint8_t* d, uint64_t len, Header* H, char* err, uint32_t err_len) {
  static const char kMarker[] = "end_header";
  const uint64_t ml = sizeof(kMarker) - 1;
  uint64_t at = UINT64_MAX;
  for (uint64_t i = 0; len >= ml && i <= len - ml; ++i)
    if (memcmp(d + i, kMarker, ml) == 0) {
      at = i;
      break;
    }
  if (at == UINT64_MAX) return fail(err, err_len, HDSR_E_INVALID_ARG, "Dropped file does not contain a complete PLY header");
  const uint64_t text_end = at + ml;
  uint64_t off = text_end;
  if (off < len && d[off] == 13) ++off;
  if (off < len && d[off] == 10) ++off;
  H->data_offset = off;

  // lines split on \r?\n
  std::vector<std::string> lines;
  std::string cur;
  for (uint64_t i = 0; i < text_end; ++i) {
    const char c = (char)d[i];
    if (c == '\n') {
      if (!cur.empty() && cur.back() == '\r') cur.pop_back();
      lines.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  lines.push_back(cur);
  if (lines[0] != "ply") return fail(err, err_len, HDSR_E_INVALID_ARG, "Dropped file does not start with a PLY header");

  bool have_format = false;
  bool in_vertex = false;
  for (size_t li = 1; li < lines.size(); ++li) {
    const std::vector<std::string> p = split_ws(lines[li]);
    if (p[0] == "format") {
      std::string f;
      for (size_t k = 1; k <= 2 && k < p.size(); ++k) f += (f.empty() ? "" : " ") + p[k];
      H->format = f;
      have_format = true;
    } else if (p[0] == "element") {
      in_vertex = p.size() > 1 && p[1] == "vertex";
      if (in_vertex) H->vertex_count = p.size() > 2 ? js_number(p[2]) : NAN;
    } else if (p[0] == "property" && in_vertex) {
      const std::string t = p.size() > 1 ? p[1] : "undefined";
      if (t == "list")
        return fail(err, err_len, HDSR_E_UNSUPPORTED, "PLY vertex list properties are not supported for splat drag-drop");
      PType pt;
      if (!normalize_type(t, &pt))
        return fail(err, err_len, HDSR_E_UNSUPPORTED, "Unsupported PLY scalar property type: " + t);
      H->props.push_back(Prop{p.size() > 2 ? p[2] : std::string(), pt, H->stride});
      H->stride += type_size(pt);
    }
  }
  const double vc = H->vertex_count;
  if (!(std::isfinite(vc) && vc == std::floor(vc) && vc > 0))
    return fail(err, err_len, HDSR_E_INVALID_ARG, "PLY vertex count must be a positive integer");
  if (H->props.empty()) return fail(err, err_len, HDSR_E_INVALID_ARG, "PLY vertex element has no scalar properties");
  const std::string fmt = have_format ? H->format : "unknown";
  // isPointCloudPlySchema (:549-558): position + rgb without splat covariance and opacity
  auto has = [&](const char* n) {
    for (const Prop& q : H->props)
      if (q.name == n) return true;
    return false;
  };
  const bool pos = has("x") && has("y") && has("z");
  const bool rgb = has("red") && has("green") && has("blue");
  const bool cov = has("scale_0") && has("scale_1") && has("scale_2") && has("rot_0") && has("rot_1") &&
                   has("rot_2") && has("rot_3");
  if (pos && rgb && !cov && !has("opacity")) {
    std::string names;
    for (size_t k = 0; k < H->props.size(); ++k) names += (k ? "," : "") + H->props[k].name;
    return fail(err, err_len, HDSR_E_UNSUPPORTED,
                "Unsupported PLY schema for hybrid splat renderer: parser=ply-point-cloud-v0; format=" + fmt +
                    "; vertexCount=" + js_int_string(vc) + "; colorSource=rgb; properties=" + names + "; file=" +
                    std::string(name ? name : "") +
                    ". This is a valid point-cloud PLY route, not a Gaussian splat PLY/SPZ payload. Kaminos cheap "
                    "preview may accept it, but the hybrid renderer needs splat covariance/opacity fields before "
                    "rendering it as splats.");
  }
  if (fmt != "binary_little_endian 1.0" && fmt != "binary_little_endian")
    return fail(err, err_len, HDSR_E_UNSUPPORTED, "Only binary_little_endian PLY drag-drop is supported");
  if (vc > 4294967295.0) return fail(err, err_len, HDSR_E_CAPACITY, "PLY vertex count exceeds 2^32 - 1");
  return HDSR_OK;
}

// new Map(properties.map(p => [p.name, p])): insertion order of first appearance, last value wins
struct FieldMap {
  std::vector<Prop> entries;
  void set(const Prop& p) {
    for (Prop& e : entries)
      if (e.name == p.name) {
        e = p;
        return;
      }
    entries.push_back(p);
  }
  const Prop* get(const char* n) const {
    for (const Prop& e : entries)
      if (e.name == n) return &e;
    return nullptr;
  }
};

struct ShLayout {
  uint32_t degree = 0, coeffs = 0;
  std::vector<const Prop*> rest;  // channel-major: [channel * coeffs + coeff]
};

bool is_rest_name(const std::string& n) {
  static const char kPre[] = "f_rest_";
  if (n.size() <= sizeof(kPre) - 1 || n.compare(0, sizeof(kPre) - 1, kPre) != 0) return false;
  for (size_t i = sizeof(kPre) - 1; i < n.size(); ++i)
    if (n[i] < '0' || n[i] > '9') return false;
  return true;
}

// discoverPlyShLayout (src/localPly.ts:429-457)
int discover_sh(const FieldMap& F, ShLayout* L, char* err, uint32_t err_len) {
  std::vector<std::pair<double, const Prop*>> rest;
  for (const Prop& e : F.entries)
    if (is_rest_name(e.name)) rest.push_back({js_number(e.name.substr(7)), &e});
  if (rest.empty()) return HDSR_OK;
  std::stable_sort(rest.begin(), rest.end(),
                   [](const std::pair<double, const Prop*>& a, const std::pair<double, const Prop*>& b) {
                     return a.first < b.first;
                   });
  if (rest.size() % 3 != 0)
    return fail(err, err_len, HDSR_E_INVALID_ARG,
                "PLY f_rest_* field count must be divisible by 3 RGB channels, got " + std::to_string(rest.size()));
  const uint32_t cc = (uint32_t)(rest.size() / 3);
  const double deg = std::sqrt((double)cc + 1.0) - 1.0;
  if (!(deg == std::floor(deg) && deg > 0))
    return fail(err, err_len, HDSR_E_INVALID_ARG,
                "PLY f_rest_* field count " + std::to_string(rest.size()) + " does not describe a complete SH degree");
  L->degree = (uint32_t)deg;
  L->coeffs = cc;
  for (const auto& r : rest) L->rest.push_back(r.second);
  return HDSR_OK;
}

void bounds_from_extents(double x0, double y0, double z0, double x1, double y1, double z1, hdsr_splat_attributes* o) {
  o->bounds_min[0] = x0;
  o->bounds_min[1] = y0;
  o->bounds_min[2] = z0;
  o->bounds_max[0] = x1;
  o->bounds_max[1] = y1;
  o->bounds_max[2] = z1;
  o->bounds_center[0] = (x0 + x1) * 0.5;
  o->bounds_center[1] = (y0 + y1) * 0.5;
  o->bounds_center[2] = (z0 + z1) * 0.5;
  o->bounds_radius = std::hypot((x1 - x0) * 0.5, (y1 - y0) * 0.5, (z1 - z0) * 0.5);
}

// TypedArray.prototype.sort order for float32: ascending, -0 before +0, NaN last
bool f32_less(float a, float b) {
  if (std::isnan(a)) return false;
  if (std::isnan(b)) return true;
  if (a < b) return true;
  if (a > b) return false;
  return std::signbit(a) && !std::signbit(b);
}

// percentileBounds (src/localPly.ts:697-721)
void percentile_bounds(const float* pos, uint32_t count, double trim, hdsr_splat_attributes* o) {
  const uint32_t m = std::min<uint32_t>(count, 50000u);
  const double step = (double)count / (double)m;
  std::vector<float> xs(m), ys(m), zs(m);
  for (uint32_t i = 0; i < m; ++i) {
    const size_t b = (size_t)std::floor((double)i * step) * 3;
    xs[i] = pos[b];
    ys[i] = pos[b + 1];
    zs[i] = pos[b + 2];
  }
  std::sort(xs.begin(), xs.end(), f32_less);
  std::sort(ys.begin(), ys.end(), f32_less);
  std::sort(zs.begin(), zs.end(), f32_less);
  const uint32_t lo = (uint32_t)std::floor((double)m * trim);
  const uint32_t hi = std::min((uint32_t)std::floor((double)m * (1.0 - trim)), m - 1);
  bounds_from_extents(xs[lo], ys[lo], zs[lo], xs[hi], ys[hi], zs[hi], o);
}

}  // namespace

extern "C" {

int hdsr_ply_inspect(const char* file_name, const void* bytes, uint64_t len, hdsr_ply_info* info, char* err,
                     uint32_t err_len) {
  if (!bytes || !info) return fail(err, err_len, HDSR_E_INVALID_ARG, "null argument");
  Header H;
  int rc = parse_header(file_name, static_cast<const uint8_t*>(bytes), len, &H, err, err_len);
  if (rc) return rc;
  FieldMap F;
  for (const Prop& p : H.props) F.set(p);
  for (const char* n : {"x", "y", "z"})
    if (!F.get(n)) return fail(err, err_len, HDSR_E_INVALID_ARG, std::string("PLY vertex property ") + n + " is required");
  ShLayout L;
  if ((rc = discover_sh(F, &L, err, err_len))) return rc;
  const uint32_t count = (uint32_t)H.vertex_count;
  if (H.data_offset + (uint64_t)count * H.stride > len)
    return fail(err, err_len, HDSR_E_INVALID_ARG,
                std::string(file_name ? file_name : "") + " ended before all " + std::to_string(count) +
                    " vertex rows could be read");
  auto has3 = [&](const char* a, const char* b, const char* c) { return F.get(a) && F.get(b) && F.get(c); };
  hdsr_ply_info I{};
  I.count = count;
  I.sh_degree = L.degree;
  I.sh_coefficient_count = L.coeffs;
  I.fields = (has3("nx", "ny", "nz") ? HDSR_PLY_NORMALS : 0u) | (F.get("roughness") ? HDSR_PLY_ROUGHNESS : 0u) |
             (F.get("metallic") ? HDSR_PLY_METALNESS : 0u) |
             (has3("emissive_r", "emissive_g", "emissive_b") ? HDSR_PLY_EMISSIVE : 0u) |
             (has3("detail_normal_x", "detail_normal_y", "detail_normal_z") ? HDSR_PLY_DETAIL_NORMALS : 0u);
  I.data_offset = H.data_offset;
  I.row_stride = H.stride;
  I.property_count = (uint32_t)H.props.size();
  *info = I;
  return HDSR_OK;
}

int hdsr_ply_decode(const char* file_name, const void* bytes, uint64_t len, hdsr_splat_attributes* o, char* err,
                    uint32_t err_len) {
  if (!bytes || !o) return fail(err, err_len, HDSR_E_INVALID_ARG, "null argument");
  hdsr_ply_info I;
  int rc = hdsr_ply_inspect(file_name, bytes, len, &I, err, err_len);
  if (rc) return rc;
  if (!o->positions || !o->colors || !o->opacities || !o->radii || !o->scales || !o->rotations || !o->original_ids ||
      (I.sh_coefficient_count && !o->sh_coefficients) || ((I.fields & HDSR_PLY_NORMALS) && !o->normals) ||
      ((I.fields & HDSR_PLY_ROUGHNESS) && !o->roughness) || ((I.fields & HDSR_PLY_METALNESS) && !o->metalness) ||
      ((I.fields & HDSR_PLY_EMISSIVE) && !o->emissive) || ((I.fields & HDSR_PLY_DETAIL_NORMALS) && !o->detail_normals))
    return fail(err, err_len, HDSR_E_INVALID_ARG, "an output array for a present attribute is null");
  Header H;
  parse_header(file_name, static_cast<const uint8_t*>(bytes), len, &H, nullptr, 0);
  FieldMap F;
  for (const Prop& p : H.props) F.set(p);
  ShLayout L;
  discover_sh(F, &L, nullptr, 0);
  auto P = [&](const char* n) { return F.get(n); };
  const Prop *px = P("x"), *py = P("y"), *pz = P("z");
  const Prop *dc0 = P("f_dc_0"), *dc1 = P("f_dc_1"), *dc2 = P("f_dc_2");
  const Prop *red = P("red"), *green = P("green"), *blue = P("blue");
  const Prop* op = P("opacity");
  const Prop *s0 = P("scale_0"), *s1 = P("scale_1"), *s2 = P("scale_2");
  const Prop *r0 = P("rot_0"), *r1 = P("rot_1"), *r2 = P("rot_2"), *r3 = P("rot_3");
  const Prop *nx = P("nx"), *ny = P("ny"), *nz = P("nz");
  const Prop *rough = P("roughness"), *metal = P("metallic");
  const Prop *er = P("emissive_r"), *eg = P("emissive_g"), *eb = P("emissive_b");
  const Prop *dx = P("detail_normal_x"), *dy = P("detail_normal_y"), *dz = P("detail_normal_z");
  const bool has_dc = dc0 && dc1 && dc2, has_rgb = red && green && blue;
  const bool has_scales = s0 && s1 && s2, has_rot = r0 && r1 && r2 && r3;
  const double kShC0 = 0.28209479177387814;
  const uint8_t* base = static_cast<const uint8_t*>(bytes) + I.data_offset;
  double mnx = INFINITY, mny = INFINITY, mnz = INFINITY, mxx = -INFINITY, mxy = -INFINITY, mxz = -INFINITY;
  const uint32_t cc = L.coeffs;
  for (uint32_t row = 0; row < I.count; ++row) {
    const uint8_t* R = base + (size_t)row * I.row_stride;
    const size_t v = (size_t)row * 3;
    const double x = read_prop(R, *px), y = read_prop(R, *py), z = read_prop(R, *pz);
    o->positions[v] = (float)x;
    o->positions[v + 1] = (float)y;
    o->positions[v + 2] = (float)z;
    mnx = js_min(mnx, x);
    mny = js_min(mny, y);
    mnz = js_min(mnz, z);
    mxx = js_max(mxx, x);
    mxy = js_max(mxy, y);
    mxz = js_max(mxz, z);
    if (has_dc) {
      o->colors[v] = (float)clamp01(0.5 + kShC0 * read_prop(R, *dc0));
      o->colors[v + 1] = (float)clamp01(0.5 + kShC0 * read_prop(R, *dc1));
      o->colors[v + 2] = (float)clamp01(0.5 + kShC0 * read_prop(R, *dc2));
    } else if (has_rgb) {
      o->colors[v] = (float)clamp01(read_prop(R, *red) / 255.0);
      o->colors[v + 1] = (float)clamp01(read_prop(R, *green) / 255.0);
      o->colors[v + 2] = (float)clamp01(read_prop(R, *blue) / 255.0);
    } else {
      o->colors[v] = o->colors[v + 1] = o->colors[v + 2] = 0.5f;
    }
    o->opacities[row] = op ? (float)(1.0 / (1.0 + std::exp(-read_prop(R, *op)))) : 1.0f;
    if (has_scales) {
      o->scales[v] = (float)read_prop(R, *s0);
      o->scales[v + 1] = (float)read_prop(R, *s1);
      o->scales[v + 2] = (float)read_prop(R, *s2);
    } else {
      o->scales[v] = o->scales[v + 1] = o->scales[v + 2] = 0.0f;
    }
    o->radii[row] = (float)js_max(js_max(std::exp((double)o->scales[v]), std::exp((double)o->scales[v + 1])),
                                  std::exp((double)o->scales[v + 2]));
    const size_t q = (size_t)row * 4;
    if (has_rot) {
      o->rotations[q] = (float)read_prop(R, *r0);
      o->rotations[q + 1] = (float)read_prop(R, *r1);
      o->rotations[q + 2] = (float)read_prop(R, *r2);
      o->rotations[q + 3] = (float)read_prop(R, *r3);
    } else {
      o->rotations[q] = 1.0f;
      o->rotations[q + 1] = o->rotations[q + 2] = o->rotations[q + 3] = 0.0f;
    }
    if (I.fields & HDSR_PLY_NORMALS) {
      o->normals[v] = (float)read_prop(R, *nx);
      o->normals[v + 1] = (float)read_prop(R, *ny);
      o->normals[v + 2] = (float)read_prop(R, *nz);
    }
    if (I.fields & HDSR_PLY_ROUGHNESS) o->roughness[row] = (float)read_prop(R, *rough);
    if (I.fields & HDSR_PLY_METALNESS) o->metalness[row] = (float)read_prop(R, *metal);
    if (I.fields & HDSR_PLY_EMISSIVE) {
      o->emissive[v] = (float)read_prop(R, *er);
      o->emissive[v + 1] = (float)read_prop(R, *eg);
      o->emissive[v + 2] = (float)read_prop(R, *eb);
    }
    if (I.fields & HDSR_PLY_DETAIL_NORMALS) {
      o->detail_normals[v] = (float)read_prop(R, *dx);
      o->detail_normals[v + 1] = (float)read_prop(R, *dy);
      o->detail_normals[v + 2] = (float)read_prop(R, *dz);
    }
    if (cc) {
      // writePlyShRow (:459-475): file order is channel-major, output coefficient-major rgb
      float* dst = o->sh_coefficients + (size_t)row * cc * 3;
      for (uint32_t k = 0; k < cc; ++k)
        for (uint32_t ch = 0; ch < 3; ++ch) dst[k * 3 + ch] = (float)read_prop(R, *L.rest[ch * cc + k]);
    }
    o->original_ids[row] = row;
  }
  if (I.count > 100)
    percentile_bounds(o->positions, I.count, 0.05, o);
  else
    bounds_from_extents(mnx, mny, mnz, mxx, mxy, mxz, o);
  return HDSR_OK;
}

}  // extern "C"
