// ply.cpp — scene ingest: the PLY header, the GPU activation of the vertex
// payload, and the quantised / exported forms of a resident scene.
//
// The header parser follows parse_header and find_end_header
// (src/scene/ply_loader.rs:234-468): binary little- or big-endian only, the
// vertex element's scalar properties in declaration order (offsets and
// stride), the 14 required float properties, f_rest_0.. counted up to the
// first gap (0 / 9 / 24 / 45 give SH degree 0-3), list properties and other
// non-empty elements rejected.  Its result is a slot table: for each of the
// 62 floats of a RawPlyVertex (raw_vertex.rs:16-37), the byte offset and
// scalar type of the property that fills it (0 = absent, left at 0), with
// a compact SH layout scattered into the 15-per-channel rest slots
// (parse_vertex_payload, 540-598).  The payload itself never goes through the
// host: it is copied to the device as is and decoded there by one kernel for
// every layout (k_ingest.hip).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "cgs_internal.h"
#include "cgs_ingest_math.h"
#include "../../include/cindergs.h"

namespace {

struct Prop {
  std::string name;
  uint32_t type;
  uint64_t offset;
};

uint32_t scalar_type(const std::string& t) {
  if (t == "char" || t == "int8") return CGS_PT_I8;
  if (t == "uchar" || t == "uint8") return CGS_PT_U8;
  if (t == "short" || t == "int16") return CGS_PT_I16;
  if (t == "ushort" || t == "uint16") return CGS_PT_U16;
  if (t == "int" || t == "int32") return CGS_PT_I32;
  if (t == "uint" || t == "uint32") return CGS_PT_U32;
  if (t == "float" || t == "float32") return CGS_PT_F32;
  if (t == "double" || t == "float64") return CGS_PT_F64;
  return CGS_PT_NONE;
}

const char* type_name(uint32_t t) {
  static const char* names[] = {"None", "I8", "U8", "I16", "U16", "I32", "U32", "F32", "F64"};
  return t <= 8u ? names[t] : "?";
}

int ply_fail(cgs_ply_layout* out, uint32_t kind, const std::string& msg, uint64_t d0 = 0,
             uint64_t d1 = 0) {
  out->error = kind;
  out->error_detail[0] = d0;
  out->error_detail[1] = d1;
  return cgs_fail(CGS_ERR_FORMAT, msg.c_str());
}

// find_end_header (234-271): the marker at a line start followed by \r\n or
// \n, else anywhere; the payload starts after the line break
bool find_end_header(const uint8_t* d, uint64_t len, uint64_t* off) {
  static const char M[] = "end_header";
  const uint64_t ml = 10;
  auto at = [&](uint64_t pos) { return pos + ml <= len && std::memcmp(d + pos, M, ml) == 0; };
  auto after = [&](uint64_t pos, uint64_t* o) {
    const uint64_t a = pos + ml;
    if (a < len) {
      if (a + 1 < len && d[a] == '\r' && d[a + 1] == '\n') {
        *o = a + 2;
        return true;
      }
      if (d[a] == '\n') {
        *o = a + 1;
        return true;
      }
    }
    return false;
  };
  for (uint64_t pos = 0; pos + ml <= len; ++pos) {
    if (!at(pos)) continue;
    const bool line_start = pos == 0 || d[pos - 1] == '\n' || d[pos - 1] == '\r';
    if (line_start && after(pos, off)) return true;
  }
  for (uint64_t pos = 0; pos + ml <= len; ++pos)
    if (at(pos)) return after(pos, off);  // the first occurrence only
  return false;
}

bool valid_utf8(const uint8_t* s, uint64_t n) {
  uint64_t i = 0;
  while (i < n) {
    const uint8_t c = s[i];
    uint32_t extra, min_cp, cp;
    if (c < 0x80) {
      ++i;
      continue;
    } else if ((c & 0xE0) == 0xC0) {
      extra = 1; min_cp = 0x80; cp = c & 0x1F;
    } else if ((c & 0xF0) == 0xE0) {
      extra = 2; min_cp = 0x800; cp = c & 0x0F;
    } else if ((c & 0xF8) == 0xF0) {
      extra = 3; min_cp = 0x10000; cp = c & 0x07;
    } else {
      return false;
    }
    if (i + extra >= n) return false;
    for (uint32_t k = 1; k <= extra; ++k) {
      if ((s[i + k] & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (s[i + k] & 0x3F);
    }
    if (cp < min_cp || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return false;
    i += extra + 1;
  }
  return true;
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

std::vector<std::string> tokens(const std::string& line) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && is_ws(line[i])) ++i;
    size_t j = i;
    while (j < line.size() && !is_ws(line[j])) ++j;
    if (j > i) out.push_back(line.substr(i, j - i));
    i = j;
  }
  return out;
}

bool parse_usize(const std::string& s, uint64_t* v) {
  if (s.empty()) return false;
  uint64_t x = 0;
  for (size_t i = (s[0] == '+' ? 1u : 0u); i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const uint64_t nx = x * 10u + (uint64_t)(s[i] - '0');
    if (nx / 10u != x) return false;
    x = nx;
  }
  *v = x;
  return s != "+";
}

const Prop* find_prop(const std::vector<Prop>& ps, const std::string& name) {
  for (const Prop& p : ps)
    if (p.name == name) return &p;
  return nullptr;
}

int parse_header(const uint8_t* bytes, uint64_t len, cgs_ply_layout* out) {
  std::memset(out, 0, sizeof(*out));
  const bool magic = (len >= 4 && std::memcmp(bytes, "ply\n", 4) == 0) ||
                     (len >= 5 && std::memcmp(bytes, "ply\r\n", 5) == 0);
  if (!magic)
    return ply_fail(out, CGS_PLY_INVALID_MAGIC,
                    "Not a valid PLY file (missing leading 'ply' magic header)");
  uint64_t payload = 0;
  if (!find_end_header(bytes, len, &payload))
    return ply_fail(out, CGS_PLY_MISSING_END_HEADER, "PLY header marker 'end_header' not found");
  if (!valid_utf8(bytes, payload))
    return ply_fail(out, CGS_PLY_INVALID_HEADER_UTF8, "PLY header is not valid UTF-8");
  const std::string text(reinterpret_cast<const char*>(bytes), payload);

  uint32_t format = 0;
  std::vector<std::pair<std::string, uint64_t>> elements;
  std::vector<Prop> props;
  uint64_t stride = 0;
  bool in_vertex = false;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t nl = text.find('\n', pos);
    if (nl == std::string::npos) nl = text.size();
    std::string line = text.substr(pos, nl - pos);
    pos = nl + 1;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    // str::trim
    size_t a = 0, b = line.size();
    while (a < b && is_ws(line[a])) ++a;
    while (b > a && is_ws(line[b - 1])) --b;
    line = line.substr(a, b - a);
    if (line.empty() || line.rfind("comment", 0) == 0 || line.rfind("obj_info", 0) == 0) continue;
    const std::vector<std::string> tk = tokens(line);
    if (tk.empty()) continue;
    if (tk[0] == "format") {
      const std::string f = tk.size() > 1 ? tk[1] : "";
      if (f == "ascii")
        return ply_fail(out, CGS_PLY_UNSUPPORTED_FORMAT,
                        "unsupported PLY format 'ascii (re-export the scene as "
                        "binary_little_endian or binary_big_endian)': only binary_little_endian "
                        "and binary_big_endian are supported");
      if (f == "binary_little_endian")
        format = CGS_PLY_BINARY_LE;
      else if (f == "binary_big_endian")
        format = CGS_PLY_BINARY_BE;
      else
        return ply_fail(out, CGS_PLY_UNSUPPORTED_FORMAT,
                        "unsupported PLY format '" + f +
                            "': only binary_little_endian and binary_big_endian are supported");
    } else if (tk[0] == "element") {
      const std::string name = tk.size() > 1 ? tk[1] : "";
      uint64_t count = 0;
      if (tk.size() < 3 || !parse_usize(tk[2], &count))
        return ply_fail(out, CGS_PLY_UNSUPPORTED_LAYOUT,
                        "unsupported vertex layout: element '" + name + "' has no count");
      if (name == "vertex")
        in_vertex = true;
      else if (in_vertex)
        in_vertex = false;
      elements.emplace_back(name, count);
    } else if (tk[0] == "property") {
      if (!in_vertex) continue;
      const std::string t = tk.size() > 1 ? tk[1] : "";
      if (t == "list") {
        const std::string name = tk.size() > 4 ? tk[4] : "property";
        return ply_fail(out, CGS_PLY_UNSUPPORTED_LIST,
                        "unsupported list property '" + name +
                            "': PLY lists are not valid in a 3DGS vertex record");
      }
      const uint32_t ty = scalar_type(t);
      if (ty == CGS_PT_NONE)
        return ply_fail(out, CGS_PLY_UNSUPPORTED_LAYOUT,
                        "unsupported vertex layout: unknown property type '" + t + "'");
      const std::string name = tk.size() > 2 ? tk[2] : "";
      if (name.empty())
        return ply_fail(out, CGS_PLY_UNSUPPORTED_LAYOUT,
                        "unsupported vertex layout: property declaration without a name");
      props.push_back({name, ty, stride});
      stride += cgs_pt_size(ty);
    }
  }
  if (format == 0) return ply_fail(out, CGS_PLY_MISSING_FORMAT, "PLY header declares no 'format' line");
  const std::pair<std::string, uint64_t>* vert = nullptr;
  for (const auto& e : elements)
    if (e.first == "vertex") {
      vert = &e;
      break;
    }
  if (!vert)
    return ply_fail(out, CGS_PLY_MISSING_VERTEX_ELEMENT,
                    "PLY header declares no 'element vertex' count");
  for (const auto& e : elements)
    if (e.first != "vertex" && e.second > 0)
      return ply_fail(out, CGS_PLY_UNSUPPORTED_ELEMENT,
                      "unsupported PLY element '" + e.first + "' with " +
                          std::to_string(e.second) + " records",
                      e.second);
  out->format = format;
  out->vertex_count = vert->second;
  out->vertex_stride = (uint32_t)stride;
  out->payload_offset = payload;
  out->num_properties = (uint32_t)props.size();
  if (vert->second == 0) return CGS_OK;
  if (props.empty())
    return ply_fail(out, CGS_PLY_UNSUPPORTED_LAYOUT,
                    "unsupported vertex layout: vertex element has no properties");
  if (stride > 0xFFFFFFF0ull)
    return ply_fail(out, CGS_PLY_UNSUPPORTED_LAYOUT, "unsupported vertex layout: record too wide");

  static const char* required[14] = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                     "scale_0", "scale_1", "scale_2", "rot_0", "rot_1",
                                     "rot_2", "rot_3"};
  for (const char* r : required) {
    const Prop* p = find_prop(props, r);
    if (!p)
      return ply_fail(out, CGS_PLY_MISSING_PROPERTY,
                      std::string("vertex property '") + r + "' is required but missing");
    if (p->type != CGS_PT_F32 && p->type != CGS_PT_F64)
      return ply_fail(out, CGS_PLY_INVALID_PROPERTY_TYPE,
                      std::string("vertex property '") + r + "' must be a float, found " +
                          type_name(p->type));
  }
  uint32_t rest = 0;
  for (;;) {
    const Prop* p = find_prop(props, "f_rest_" + std::to_string(rest));
    if (!p) break;
    if (p->type != CGS_PT_F32 && p->type != CGS_PT_F64)
      return ply_fail(out, CGS_PLY_INVALID_PROPERTY_TYPE,
                      "vertex property '" + p->name + "' must be a float, found " +
                          type_name(p->type));
    ++rest;
  }
  const int deg = rest == 0 ? 0 : rest == 9 ? 1 : rest == 24 ? 2 : rest == 45 ? 3 : -1;
  if (deg < 0)
    return ply_fail(out, CGS_PLY_UNSUPPORTED_SH_LAYOUT,
                    "spherical harmonics layout with " + std::to_string(rest) +
                        " f_rest properties is unsupported (expected 0, 9, 24, or 45)",
                    rest);
  out->sh_degree = (uint32_t)deg;

  // slot table of the 62 RawPlyVertex floats
  auto set = [&](uint32_t slot, const std::string& name) {
    const Prop* p = find_prop(props, name);
    if (p) {
      out->slot_offset[slot] = (uint32_t)p->offset;
      out->slot_type[slot] = (uint8_t)p->type;
    }
  };
  set(0, "x"); set(1, "y"); set(2, "z");
  set(3, "nx"); set(4, "ny"); set(5, "nz");
  for (uint32_t c = 0; c < 3; ++c) set(6 + c, "f_dc_" + std::to_string(c));
  const uint32_t per_ch = rest / 3u;
  for (uint32_t c = 0; c < 3; ++c)
    for (uint32_t i = 0; i < per_ch; ++i)
      set(9 + c * 15 + i, "f_rest_" + std::to_string(c * per_ch + i));
  set(54, "opacity");
  for (uint32_t c = 0; c < 3; ++c) set(55 + c, "scale_" + std::to_string(c));
  for (uint32_t c = 0; c < 4; ++c) set(58 + c, "rot_" + std::to_string(c));

  // is_canonical_layout (491-516): the 62 names in RawPlyVertex order, all f32, LE
  bool canonical = format == CGS_PLY_BINARY_LE && stride == 248 && props.size() == 62;
  for (uint32_t k = 0; canonical && k < 62; ++k) {
    static const char* head[6] = {"x", "y", "z", "nx", "ny", "nz"};
    std::string want;
    if (k < 6) want = head[k];
    else if (k < 9) want = "f_dc_" + std::to_string(k - 6);
    else if (k < 54) want = "f_rest_" + std::to_string(k - 9);
    else if (k == 54) want = "opacity";
    else if (k < 58) want = "scale_" + std::to_string(k - 55);
    else want = "rot_" + std::to_string(k - 58);
    canonical = props[k].name == want && props[k].type == CGS_PT_F32;
  }
  out->canonical = canonical ? 1u : 0u;
  return CGS_OK;
}

// kernel arguments of a layout: every present slot inside the record, the
// word path when all present slots are aligned little-endian float32
int activate_args(const cgs_ply_layout* L, CgsActivateArgs* A) {
  if (L->vertex_stride == 0 || L->vertex_stride > 65528u)
    return cgs_fail(CGS_ERR_INVALID_ARG, "vertex_stride must be in [1, 65528] bytes");
  if (L->format != CGS_PLY_BINARY_LE && L->format != CGS_PLY_BINARY_BE)
    return cgs_fail(CGS_ERR_INVALID_ARG, "format must be binary little- or big-endian");
  bool words = L->format == CGS_PLY_BINARY_LE && L->vertex_stride % 4u == 0u;
  for (uint32_t k = 0; k < 62u; ++k) {
    const uint32_t t = L->slot_type[k];
    if (t > CGS_PT_F64) return cgs_fail(CGS_ERR_INVALID_ARG, "unknown slot type");
    if (t == CGS_PT_NONE) continue;
    if ((uint64_t)L->slot_offset[k] + cgs_pt_size(t) > L->vertex_stride)
      return cgs_fail(CGS_ERR_INVALID_ARG, "a slot lies outside the vertex record");
    words = words && t == CGS_PT_F32 && L->slot_offset[k] % 4u == 0u;
  }
  A->stride = L->vertex_stride;
  A->le = L->format == CGS_PLY_BINARY_LE ? 1u : 0u;
  A->per_block = (uint32_t)std::min<uint64_t>(128u, (65536u - 8u) / L->vertex_stride);
  A->word_path = words ? 1u : 0u;
  std::memcpy(A->slot_offset, L->slot_offset, sizeof(A->slot_offset));
  std::memcpy(A->slot_type, L->slot_type, sizeof(A->slot_type));
  return CGS_OK;
}

bool misaligned(const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) != 0u; }

int finish(cgs_context* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(cgs_context_stream(c));
  if (e != hipSuccess) return cgs_fail(CGS_ERR_EXECUTION, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
  return CGS_OK;
}

}  // namespace

extern "C" {

int cgs_ply_parse_header(const uint8_t* bytes, uint64_t len, cgs_ply_layout* out) {
  if (!out) return cgs_fail(CGS_ERR_INVALID_ARG, "null layout");
  if (!bytes && len) return cgs_fail(CGS_ERR_INVALID_ARG, "null bytes");
  return parse_header(bytes ? bytes : reinterpret_cast<const uint8_t*>(""), len, out);
}

int cgs_activate_raw(cgs_context* c, const void* d_records, const cgs_ply_layout* layout,
                     uint64_t n, float* d_pos, float* d_cov, float* d_opa, float* d_sh) {
  const int rc = cgs_context_enter(c);
  if (rc != CGS_OK) return rc;
  if (!layout) return cgs_fail(CGS_ERR_INVALID_ARG, "null layout");
  if (n == 0) return CGS_OK;
  if (!d_records || !d_pos || !d_cov || !d_opa || !d_sh)
    return cgs_fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (misaligned(d_records, 4) || misaligned(d_cov, 8) || misaligned(d_sh, 16))
    return cgs_fail(CGS_ERR_INVALID_ARG,
                    "d_records must be 4-byte, d_cov 8-byte and d_sh 16-byte aligned");
  if (n > (1ull << 40) / (layout->vertex_stride ? layout->vertex_stride : 1u))
    return cgs_fail(CGS_ERR_INVALID_ARG, "payload too large");
  CgsActivateArgs A;
  const int ra = activate_args(layout, &A);
  if (ra != CGS_OK) return ra;
  A.records = static_cast<const uint8_t*>(d_records);
  A.n = n;
  A.pos = d_pos;
  A.cov = d_cov;
  A.opa = d_opa;
  A.sh = d_sh;
  if ((n + A.per_block - 1u) / A.per_block > 0x7FFFFFFFull)
    return cgs_fail(CGS_ERR_INVALID_ARG, "too many records");
  cgs_context_prof_begin(c);
  cgs_context_prof_mark(c, "k_activate_raw");
  cgs_launch_activate_raw(A, cgs_context_stream(c));
  const int r = finish(c, "activate raw");
  cgs_context_prof_end(c, r == CGS_OK);
  return r;
}

int cgs_ply_load(cgs_context* c, const uint8_t* bytes, uint64_t len, const cgs_ply_layout* layout,
                 float* d_pos, float* d_cov, float* d_opa, float* d_sh) {
  int rc = cgs_context_enter(c);
  if (rc != CGS_OK) return rc;
  cgs_ply_layout parsed;
  if (!layout) {
    rc = cgs_ply_parse_header(bytes, len, &parsed);
    if (rc != CGS_OK) return rc;
    layout = &parsed;
  }
  const uint64_t n = layout->vertex_count;
  if (n == 0) return CGS_OK;
  if (!bytes || layout->payload_offset > len) return cgs_fail(CGS_ERR_INVALID_ARG, "bad bytes");
  const uint64_t avail = len - layout->payload_offset;
  // n * stride can wrap for a huge vertex count: compare by division first
  const uint64_t stride = layout->vertex_stride;
  const bool too_many = stride != 0 && n > avail / stride;
  const uint64_t needed = too_many ? UINT64_MAX : n * stride;
  if (too_many || avail < needed) {
    cgs_ply_layout err = *layout;
    return ply_fail(&err, CGS_PLY_TRUNCATED_PAYLOAD,
                    "vertex payload is truncated: need " + std::to_string(needed) +
                        " bytes, file provides " + std::to_string(avail),
                    needed, avail);
  }
  void* stage = nullptr;
  rc = cgs_context_staging(c, needed, &stage);
  if (rc != CGS_OK) return rc;
  const hipError_t e = hipMemcpyAsync(stage, bytes + layout->payload_offset, needed,
                                      hipMemcpyHostToDevice, cgs_context_stream(c));
  if (e != hipSuccess)
    return cgs_fail(CGS_ERR_EXECUTION, (std::string("payload upload: ") + hipGetErrorString(e)).c_str());
  return cgs_activate_raw(c, stage, layout, n, d_pos, d_cov, d_opa, d_sh);
}

int cgs_quantize_scene(cgs_context* c, const float* d_cov, const float* d_sh, uint64_t n,
                       uint32_t sh_degree, uint32_t* d_cov_f16, uint32_t* d_rot_oct,
                       uint32_t* d_sh_f16) {
  const int rc = cgs_context_enter(c);
  if (rc != CGS_OK) return rc;
  if (n == 0) return CGS_OK;
  if (!d_cov || !d_sh || !d_cov_f16 || !d_rot_oct || !d_sh_f16)
    return cgs_fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (misaligned(d_sh, 16) || misaligned(d_cov, 8))
    return cgs_fail(CGS_ERR_INVALID_ARG, "d_sh must be 16-byte and d_cov 8-byte aligned");
  if (n >= (1ull << 31) * 256ull) return cgs_fail(CGS_ERR_INVALID_ARG, "n too large");
  cgs_context_prof_begin(c);
  cgs_context_prof_mark(c, "k_quantize");
  cgs_launch_quantize(d_cov, d_sh, n, sh_degree > 3u ? 3u : sh_degree, d_cov_f16, d_rot_oct,
                      d_sh_f16, cgs_context_stream(c));
  const int r = finish(c, "quantize");
  cgs_context_prof_end(c, r == CGS_OK);
  return r;
}

int cgs_export_raw(cgs_context* c, const float* d_pos, const float* d_cov, const float* d_opa,
                   const float* d_sh, uint64_t n, void* d_records) {
  const int rc = cgs_context_enter(c);
  if (rc != CGS_OK) return rc;
  if (n == 0) return CGS_OK;
  if (!d_pos || !d_cov || !d_opa || !d_sh || !d_records)
    return cgs_fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (misaligned(d_sh, 16) || misaligned(d_records, 8))
    return cgs_fail(CGS_ERR_INVALID_ARG, "d_sh must be 16-byte and d_records 8-byte aligned");
  if (n >= (1ull << 31) * 256ull) return cgs_fail(CGS_ERR_INVALID_ARG, "n too large");
  cgs_context_prof_begin(c);
  cgs_context_prof_mark(c, "k_export_raw");
  cgs_launch_export_raw(d_pos, d_cov, d_opa, d_sh, n, static_cast<float*>(d_records),
                        cgs_context_stream(c));
  const int r = finish(c, "export raw");
  cgs_context_prof_end(c, r == CGS_OK);
  return r;
}

}  // extern "C"
