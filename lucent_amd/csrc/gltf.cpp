// glTF 2.0 / GLB ingestion behind the C ABI (lucent_gltf_*): the behaviour of
// the reference's loadGLTF (src/loaders/gltf.ts:188-502) — world-space
// triangle soups per primitive, its material model, only the images some
// material samples, returned encoded — in a two-phase interface that suits a
// synchronous C caller: lucent_gltf_parse reads the container and lists the
// external resources the asset needs (buffers, then the images materials
// use, in first-use order), the caller supplies their bytes, and
// lucent_gltf_build assembles meshes, materials and images.
//
// Numerics follow the reference: accessors decode to float32, node
// transforms compose in float64, world positions and normals are computed in
// float64 and stored as float32.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

namespace lucent {
namespace {

// ---------------------------------------------------------------------------
// A small JSON reader: values as a tree, objects keep their key order.
// ---------------------------------------------------------------------------
struct Json {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* get(const char* key) const {
    if (kind != kObj) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  const Json* at(size_t i) const { return kind == kArr && i < arr.size() ? &arr[i] : nullptr; }
  bool is_num() const { return kind == kNum; }
};

double num_or(const Json* j, double d) { return j && j->is_num() ? j->num : d; }
// glTF indices are JSON numbers; -1 when absent.
long idx_or(const Json* j) { return j && j->is_num() ? long(j->num) : -1; }

class JsonReader {
 public:
  JsonReader(const char* p, size_t n) : p_(p), end_(p + n) {}
  Status read(Json* out) {
    skip();
    Status s = value(out, 0);
    if (s != LUCENT_OK) return s;
    skip();
    if (p_ != end_) return bad("trailing characters after the JSON value");
    return LUCENT_OK;
  }

 private:
  const char* p_;
  const char* end_;

  static Status bad(const std::string& what) { return fail(LUCENT_E_INVALID, "glTF JSON: " + what); }
  void skip() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (size_t(end_ - p_) < n || std::memcmp(p_, w, n) != 0) return false;
    p_ += n;
    return true;
  }
  Status value(Json* v, int depth) {
    if (depth > 256) return bad("nesting too deep");
    if (p_ >= end_) return bad("unexpected end");
    switch (*p_) {
      case '{': return object(v, depth);
      case '[': return array(v, depth);
      case '"': v->kind = Json::kStr; return string(&v->str);
      case 't': if (!lit("true")) return bad("bad literal"); v->kind = Json::kBool; v->b = true; return LUCENT_OK;
      case 'f': if (!lit("false")) return bad("bad literal"); v->kind = Json::kBool; v->b = false; return LUCENT_OK;
      case 'n': if (!lit("null")) return bad("bad literal"); v->kind = Json::kNull; return LUCENT_OK;
      default: return number(v);
    }
  }
  Status object(Json* v, int depth) {
    v->kind = Json::kObj;
    ++p_;
    skip();
    if (p_ < end_ && *p_ == '}') {
      ++p_;
      return LUCENT_OK;
    }
    for (;;) {
      skip();
      if (p_ >= end_ || *p_ != '"') return bad("expected a key");
      std::string key;
      Status s = string(&key);
      if (s != LUCENT_OK) return s;
      skip();
      if (p_ >= end_ || *p_ != ':') return bad("expected ':'");
      ++p_;
      skip();
      v->obj.emplace_back(std::move(key), Json());
      if ((s = value(&v->obj.back().second, depth + 1)) != LUCENT_OK) return s;
      skip();
      if (p_ < end_ && *p_ == ',') {
        ++p_;
        continue;
      }
      if (p_ < end_ && *p_ == '}') {
        ++p_;
        return LUCENT_OK;
      }
      return bad("expected ',' or '}'");
    }
  }
  Status array(Json* v, int depth) {
    v->kind = Json::kArr;
    ++p_;
    skip();
    if (p_ < end_ && *p_ == ']') {
      ++p_;
      return LUCENT_OK;
    }
    for (;;) {
      skip();
      v->arr.emplace_back();
      Status s = value(&v->arr.back(), depth + 1);
      if (s != LUCENT_OK) return s;
      skip();
      if (p_ < end_ && *p_ == ',') {
        ++p_;
        continue;
      }
      if (p_ < end_ && *p_ == ']') {
        ++p_;
        return LUCENT_OK;
      }
      return bad("expected ',' or ']'");
    }
  }
  static void put_utf8(std::string* s, uint32_t c) {
    if (c < 0x80) {
      s->push_back(char(c));
    } else if (c < 0x800) {
      s->push_back(char(0xC0 | (c >> 6)));
      s->push_back(char(0x80 | (c & 0x3F)));
    } else if (c < 0x10000) {
      s->push_back(char(0xE0 | (c >> 12)));
      s->push_back(char(0x80 | ((c >> 6) & 0x3F)));
      s->push_back(char(0x80 | (c & 0x3F)));
    } else {
      s->push_back(char(0xF0 | (c >> 18)));
      s->push_back(char(0x80 | ((c >> 12) & 0x3F)));
      s->push_back(char(0x80 | ((c >> 6) & 0x3F)));
      s->push_back(char(0x80 | (c & 0x3F)));
    }
  }
  Status hex4(uint32_t* out) {
    if (end_ - p_ < 4) return bad("short \\u escape");
    uint32_t c = 0;
    for (int i = 0; i < 4; ++i) {
      const char h = *p_++;
      c <<= 4;
      if (h >= '0' && h <= '9') c |= uint32_t(h - '0');
      else if (h >= 'a' && h <= 'f') c |= uint32_t(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') c |= uint32_t(h - 'A' + 10);
      else return bad("bad \\u escape");
    }
    *out = c;
    return LUCENT_OK;
  }
  Status string(std::string* out) {
    ++p_;  // opening quote
    while (p_ < end_ && *p_ != '"') {
      if (*p_ != '\\') {
        out->push_back(*p_++);
        continue;
      }
      if (++p_ >= end_) return bad("unterminated escape");
      const char e = *p_++;
      switch (e) {
        case '"': out->push_back('"'); break;
        case '\\': out->push_back('\\'); break;
        case '/': out->push_back('/'); break;
        case 'b': out->push_back('\b'); break;
        case 'f': out->push_back('\f'); break;
        case 'n': out->push_back('\n'); break;
        case 'r': out->push_back('\r'); break;
        case 't': out->push_back('\t'); break;
        case 'u': {
          uint32_t c = 0;
          Status s = hex4(&c);
          if (s != LUCENT_OK) return s;
          if (c >= 0xD800 && c < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
            p_ += 2;
            uint32_t lo = 0;
            if ((s = hex4(&lo)) != LUCENT_OK) return s;
            c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(out, c);
          break;
        }
        default: return bad("bad escape");
      }
    }
    if (p_ >= end_) return bad("unterminated string");
    ++p_;
    return LUCENT_OK;
  }
  Status number(Json* v) {
    const char* s = p_;
    if (p_ < end_ && (*p_ == '-' || *p_ == '+')) ++p_;
    while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '-' ||
                         *p_ == '+'))
      ++p_;
    if (p_ == s) return bad(std::string("unexpected character '") + *s + "'");
    const std::string text(s, p_);
    char* stop = nullptr;
    v->num = std::strtod(text.c_str(), &stop);
    if (!stop || *stop != '\0') return bad("bad number " + text);
    v->kind = Json::kNum;
    return LUCENT_OK;
  }
};

// ---------------------------------------------------------------------------
// Bytes: GLB container, data URIs.
// ---------------------------------------------------------------------------
constexpr uint32_t kGlbMagic = 0x46546C67, kChunkJson = 0x4E4F534A, kChunkBin = 0x004E4942;

uint32_t le32(const uint8_t* p) { return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24; }

int b64(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+' || c == '-') return 62;
  if (c == '/' || c == '_') return 63;
  return -1;
}

int hexval(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}

// data:[<mime>][;base64],<payload>: base64 or percent-encoded.
Status decode_data_uri(const std::string& uri, std::vector<uint8_t>* out) {
  const size_t comma = uri.find(',');
  if (comma == std::string::npos) return fail(LUCENT_E_INVALID, "malformed data URI");
  const std::string meta = uri.substr(5, comma - 5);
  const bool base64 = meta.size() >= 7 && meta.compare(meta.size() - 7, 7, ";base64") == 0;
  out->clear();
  if (!base64) {
    for (size_t i = comma + 1; i < uri.size(); ++i) {
      if (uri[i] == '%' && i + 2 < uri.size() && hexval(uri[i + 1]) >= 0 && hexval(uri[i + 2]) >= 0) {
        out->push_back(uint8_t(hexval(uri[i + 1]) * 16 + hexval(uri[i + 2])));
        i += 2;
      } else {
        out->push_back(uint8_t(uri[i]));
      }
    }
    return LUCENT_OK;
  }
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = comma + 1; i < uri.size(); ++i) {
    const char c = uri[i];
    if (c == '=' || c == ' ' || c == '\n' || c == '\r') continue;
    const int v = b64(c);
    if (v < 0) return fail(LUCENT_E_INVALID, "malformed base64 in data URI");
    acc = acc << 6 | uint32_t(v);
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out->push_back(uint8_t(acc >> bits));
    }
  }
  return LUCENT_OK;
}

std::string mime_of_data_uri(const std::string& uri) {
  size_t e = 5;
  while (e < uri.size() && uri[e] != ';' && uri[e] != ',') ++e;
  return uri.substr(5, e - 5);
}

std::string mime_of_name(const std::string& name) {
  std::string base = name.substr(0, name.find_first_of("?#"));
  const size_t dot = base.rfind('.');
  std::string ext = dot == std::string::npos ? "" : base.substr(dot + 1);
  for (auto& c : ext) c = char(std::tolower(static_cast<unsigned char>(c)));
  if (ext == "png") return "image/png";
  if (ext == "jpg" || ext == "jpeg") return "image/jpeg";
  if (ext == "webp") return "image/webp";
  if (ext == "ktx2") return "image/ktx2";
  return "application/octet-stream";
}

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

// ---------------------------------------------------------------------------
// The document.
// ---------------------------------------------------------------------------
struct Resource {
  std::string uri;
  bool ready = false;
  std::vector<uint8_t> bytes;
};

struct Image {
  long source = -1;  // index into json images
  std::vector<uint8_t> bytes;
  std::string mime, name;
  long resource = -1;  // external resource it waits for
};

struct Mesh {
  std::vector<float> pos, nrm, uv;
  int32_t material = 0;
};

}  // namespace
}  // namespace lucent

struct lucent_gltf {
  lucent::Json json;
  std::vector<uint8_t> glb_bin;
  bool has_glb_bin = false;
  std::vector<lucent::Resource> resources;
  // Per JSON buffer: its bytes (GLB chunk or data URI) or the resource it waits for.
  std::vector<std::vector<uint8_t>> buffer_bytes;
  std::vector<long> buffer_resource;
  std::vector<lucent::Image> images;   // the images materials sample, in first-use order
  std::map<long, int32_t> image_slot;  // JSON image index -> slot
  std::vector<std::string> warnings;
  // Output of lucent_gltf_build.
  bool built = false;
  std::vector<lucent::Mesh> meshes;
  std::vector<std::array<double, 40>> materials;  // material texel layout, unrounded
  double bounds[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  uint64_t triangles = 0;
  std::string error_scratch;
};

namespace lucent {
namespace {

using Doc = lucent_gltf;

void warn(Doc* d, const std::string& w) {
  if (std::find(d->warnings.begin(), d->warnings.end(), w) == d->warnings.end()) d->warnings.push_back(w);
}

const Json* jget(const Json* j, const char* key) { return j ? j->get(key) : nullptr; }

// Extensions whose absence breaks geometry or textures (refused), and the
// material extensions the material model covers.
const char* const kRefusedRequired[] = {"KHR_draco_mesh_compression", "EXT_meshopt_compression", "KHR_texture_basisu",
                                        "MSFT_texture_dds", "EXT_mesh_gpu_instancing"};
const char* const kMaterialExtensions[] = {"KHR_materials_emissive_strength", "KHR_materials_transmission",
                                           "KHR_materials_ior", "KHR_materials_volume", "KHR_materials_clearcoat",
                                           "KHR_materials_sheen", "KHR_materials_unlit"};

long add_resource(Doc* d, const std::string& uri) {
  d->resources.push_back(Resource{uri, false, {}});
  return long(d->resources.size() - 1);
}

// The image slot a texture reference samples (registering the image on first
// use), or -1 (textureSlot's rules: TEXCOORD_0 only, WebP source, no KTX2).
int32_t texture_slot(Doc* d, const Json* info) {
  const Json* index = jget(info, "index");
  if (!index || !index->is_num()) return -1;
  const Json* textures = d->json.get("textures");
  const Json* t = textures ? textures->at(size_t(index->num)) : nullptr;
  if (!t) return -1;
  if (num_or(jget(info, "texCoord"), 0.0) != 0.0) warn(d, "textures on TEXCOORD_1 or higher are sampled with TEXCOORD_0");
  const Json* ext = jget(t, "extensions");
  long source = idx_or(jget(jget(ext, "EXT_texture_webp"), "source"));
  if (source < 0) source = idx_or(jget(t, "source"));
  if (source < 0) {
    if (jget(ext, "KHR_texture_basisu")) warn(d, "KTX2/Basis textures are not supported and were skipped");
    return -1;
  }
  const Json* imgs = d->json.get("images");
  const Json* img = imgs ? imgs->at(size_t(source)) : nullptr;
  const Json* mime_j = jget(img, "mimeType");
  const Json* uri_j = jget(img, "uri");
  const std::string mime = mime_j && mime_j->kind == Json::kStr ? mime_j->str
                                                                : mime_of_name(uri_j && uri_j->kind == Json::kStr ? uri_j->str : "");
  if (mime == "image/ktx2") {
    warn(d, "KTX2/Basis textures are not supported and were skipped");
    return -1;
  }
  auto it = d->image_slot.find(source);
  if (it != d->image_slot.end()) return it->second;
  if (!img) return -1;
  Image im;
  im.source = source;
  const Json* name_j = jget(img, "name");
  im.name = name_j && name_j->kind == Json::kStr ? name_j->str
            : uri_j && uri_j->kind == Json::kStr ? uri_j->str
                                                 : "image " + std::to_string(source);
  if (jget(img, "bufferView")) {
    im.mime = mime_j && mime_j->kind == Json::kStr ? mime_j->str : "image/png";
  } else if (uri_j && uri_j->kind == Json::kStr && starts_with(uri_j->str, "data:")) {
    im.mime = mime_of_data_uri(uri_j->str);
  } else if (uri_j && uri_j->kind == Json::kStr) {
    im.mime = mime_j && mime_j->kind == Json::kStr ? mime_j->str : mime_of_name(uri_j->str);
    im.resource = add_resource(d, uri_j->str);
  }
  const int32_t slot = int32_t(d->images.size());
  d->images.push_back(std::move(im));
  d->image_slot[source] = slot;
  return slot;
}

// A bufferView's bytes and stride.
Status view_bytes(const Doc* d, long view, const uint8_t** p, size_t* len, size_t* stride) {
  const Json* views = d->json.get("bufferViews");
  const Json* v = views ? views->at(size_t(view)) : nullptr;
  if (!v || view < 0) return fail(LUCENT_E_INVALID, "missing bufferView " + std::to_string(view));
  const long b = idx_or(v->get("buffer"));
  if (b < 0 || size_t(b) >= d->buffer_bytes.size())
    return fail(LUCENT_E_INVALID, "bufferView " + std::to_string(view) + " references missing buffer " + std::to_string(b));
  const std::vector<uint8_t>& buf =
      d->buffer_resource[size_t(b)] >= 0 ? d->resources[size_t(d->buffer_resource[size_t(b)])].bytes : d->buffer_bytes[size_t(b)];
  const double off = num_or(v->get("byteOffset"), 0.0), n = num_or(v->get("byteLength"), 0.0);
  if (off < 0 || n < 0 || off + n > double(buf.size()))
    return fail(LUCENT_E_INVALID, "bufferView " + std::to_string(view) + " runs past its buffer");
  *p = buf.data() + size_t(off);
  *len = size_t(n);
  *stride = size_t(num_or(v->get("byteStride"), 0.0));
  return LUCENT_OK;
}

int components_of(const std::string& type) {
  if (type == "SCALAR") return 1;
  if (type == "VEC2") return 2;
  if (type == "VEC3") return 3;
  if (type == "VEC4" || type == "MAT2") return 4;
  if (type == "MAT3") return 9;
  if (type == "MAT4") return 16;
  return 0;
}
int bytes_of(long component_type) {
  switch (component_type) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: return 0;
  }
}
// One component as float32, with the accessor's normalisation.
float read_component(const uint8_t* p, long type, bool normalized) {
  switch (type) {
    case 5126: { float f; std::memcpy(&f, p, 4); return f; }
    case 5125: return float(double(le32(p)));
    case 5123: { const double v = double(uint16_t(p[0] | p[1] << 8)); return float(normalized ? v / 65535.0 : v); }
    case 5122: { const double v = double(int16_t(uint16_t(p[0] | p[1] << 8))); return float(normalized ? std::max(v / 32767.0, -1.0) : v); }
    case 5121: { const double v = double(p[0]); return float(normalized ? v / 255.0 : v); }
    default: { const double v = double(int8_t(p[0])); return float(normalized ? std::max(v / 127.0, -1.0) : v); }
  }
}

struct Accessor {
  std::vector<float> data;
  int size = 0;
  size_t count = 0;
};

Status read_accessor(Doc* d, long index, std::map<long, Accessor>* cache, const Accessor** out) {
  auto it = cache->find(index);
  if (it != cache->end()) {
    *out = &it->second;
    return LUCENT_OK;
  }
  const Json* accs = d->json.get("accessors");
  const Json* a = accs ? accs->at(size_t(index)) : nullptr;
  if (!a || index < 0) return fail(LUCENT_E_INVALID, "missing accessor " + std::to_string(index));
  const Json* type_j = a->get("type");
  const long ctype = idx_or(a->get("componentType"));
  const int size = components_of(type_j && type_j->kind == Json::kStr ? type_j->str : "");
  const int cbytes = bytes_of(ctype);
  if (!size || !cbytes)
    return fail(LUCENT_E_INVALID, "accessor " + std::to_string(index) + " has unsupported type " +
                                      (type_j && type_j->kind == Json::kStr ? type_j->str : "?") + "/" + std::to_string(ctype));
  const Json* norm_j = a->get("normalized");
  const bool normalized = norm_j && norm_j->kind == Json::kBool && norm_j->b;
  Accessor acc;
  acc.size = size;
  acc.count = size_t(num_or(a->get("count"), 0.0));
  acc.data.assign(acc.count * size, 0.0f);
  const long view = idx_or(a->get("bufferView"));
  if (view >= 0) {
    const uint8_t* p;
    size_t len, stride;
    Status s = view_bytes(d, view, &p, &len, &stride);
    if (s != LUCENT_OK) return s;
    const size_t elem = size_t(size) * cbytes, step = stride ? stride : elem;
    const size_t base = size_t(num_or(a->get("byteOffset"), 0.0));
    if (acc.count > 0 && base + step * (acc.count - 1) + elem > len)
      return fail(LUCENT_E_INVALID, "accessor " + std::to_string(index) + " reads past its bufferView");
    for (size_t i = 0; i < acc.count; ++i)
      for (int c = 0; c < size; ++c) acc.data[i * size + c] = read_component(p + base + i * step + size_t(c) * cbytes, ctype, normalized);
  }
  if (const Json* sp = a->get("sparse")) {
    const size_t n = size_t(num_or(sp->get("count"), 0.0));
    const Json* ij = sp->get("indices");
    const Json* vj = sp->get("values");
    const uint8_t *ip, *vp;
    size_t il, vl, is, vs;
    Status s = view_bytes(d, idx_or(jget(ij, "bufferView")), &ip, &il, &is);
    if (s != LUCENT_OK) return s;
    if ((s = view_bytes(d, idx_or(jget(vj, "bufferView")), &vp, &vl, &vs)) != LUCENT_OK) return s;
    const long itype = idx_or(jget(ij, "componentType"));
    const int ib = bytes_of(itype);
    const size_t ioff = size_t(num_or(jget(ij, "byteOffset"), 0.0)), voff = size_t(num_or(jget(vj, "byteOffset"), 0.0));
    if (!ib || ioff + n * ib > il || voff + n * size * cbytes > vl)
      return fail(LUCENT_E_INVALID, "sparse data of accessor " + std::to_string(index) + " runs past its bufferView");
    for (size_t k = 0; k < n; ++k) {
      const size_t target = size_t(read_component(ip + ioff + k * ib, itype, false));
      if (target >= acc.count)
        return fail(LUCENT_E_INVALID, "sparse index " + std::to_string(target) + " out of range in accessor " + std::to_string(index));
      for (int c = 0; c < size; ++c)
        acc.data[target * size + c] = read_component(vp + voff + (k * size + c) * cbytes, ctype, normalized);
    }
  }
  auto ins = cache->emplace(index, std::move(acc));
  *out = &ins.first->second;
  return LUCENT_OK;
}

using Mat4 = std::array<double, 16>;  // column-major

Mat4 identity4() {
  Mat4 m{};
  m[0] = m[5] = m[10] = m[15] = 1.0;
  return m;
}
Mat4 mul4(const Mat4& a, const Mat4& b) {
  Mat4 o{};
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
      o[c * 4 + r] = s;
    }
  return o;
}
// T * R * S of a node (column-major), the quaternion as (x, y, z, w).
Mat4 trs(const double t[3], const double q[4], const double s[3]) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  Mat4 m{};
  m[0] = (1 - 2 * (y * y + z * z)) * s[0];
  m[1] = 2 * (x * y + z * w) * s[0];
  m[2] = 2 * (x * z - y * w) * s[0];
  m[4] = 2 * (x * y - z * w) * s[1];
  m[5] = (1 - 2 * (x * x + z * z)) * s[1];
  m[6] = 2 * (y * z + x * w) * s[1];
  m[8] = 2 * (x * z + y * w) * s[2];
  m[9] = 2 * (y * z - x * w) * s[2];
  m[10] = (1 - 2 * (x * x + y * y)) * s[2];
  m[12] = t[0];
  m[13] = t[1];
  m[14] = t[2];
  m[15] = 1;
  return m;
}

void read_vec(const Json* j, double* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = num_or(j ? j->at(size_t(i)) : nullptr, out[i]);
}

// The material model of material.ts (packed as its 10 texels) from a glTF
// material (pbrMetallicRoughness, emissive, normal, alpha, and the KHR
// material extensions).
std::array<double, 40> convert_material(Doc* d, const Json* m) {
  const Json* pbr = jget(m, "pbrMetallicRoughness");
  const Json* ext = jget(m, "extensions");
  if (ext && ext->kind == Json::kObj)
    for (const auto& kv : ext->obj) {
      bool known = false;
      for (const char* h : kMaterialExtensions) known = known || kv.first == h;
      if (!known) warn(d, "material extension " + kv.first + " is ignored");
    }
  double bc[4] = {1, 1, 1, 1};
  read_vec(jget(pbr, "baseColorFactor"), bc, 4);
  double ef[3] = {0, 0, 0};
  read_vec(jget(m, "emissiveFactor"), ef, 3);
  const double strength = num_or(jget(jget(ext, "KHR_materials_emissive_strength"), "emissiveStrength"), 1.0);
  const Json* alpha_j = jget(m, "alphaMode");
  const std::string alpha_mode = alpha_j && alpha_j->kind == Json::kStr ? alpha_j->str : "OPAQUE";
  const float alpha_code = alpha_mode == "MASK" ? 1.0f : alpha_mode == "BLEND" ? 2.0f : 0.0f;
  const Json* infos[4] = {jget(pbr, "baseColorTexture"), jget(pbr, "metallicRoughnessTexture"), jget(m, "normalTexture"),
                          jget(m, "emissiveTexture")};
  const int32_t base_map = texture_slot(d, infos[0]);
  const int32_t mr_map = texture_slot(d, infos[1]);
  const int32_t normal_map = texture_slot(d, infos[2]);
  const int32_t emissive_map = texture_slot(d, infos[3]);
  // KHR_texture_transform of the first map that has one: T * R * S rows.
  double uv[6] = {1, 0, 0, 0, 1, 0};
  for (const Json* info : infos) {
    const Json* tt = jget(jget(info, "extensions"), "KHR_texture_transform");
    if (!tt) continue;
    double off[2] = {0, 0}, sc[2] = {1, 1};
    read_vec(tt->get("offset"), off, 2);
    read_vec(tt->get("scale"), sc, 2);
    const double r = num_or(tt->get("rotation"), 0.0), c = std::cos(r), s = std::sin(r);
    const double t[6] = {c * sc[0], -s * sc[1], off[0], s * sc[0], c * sc[1], off[1]};
    std::copy(t, t + 6, uv);
    break;
  }
  const Json* volume = jget(ext, "KHR_materials_volume");
  double att_color[3] = {1, 1, 1};
  read_vec(jget(volume, "attenuationColor"), att_color, 3);
  const double att_dist = num_or(jget(volume, "attenuationDistance"), INFINITY);
  const bool thin = !volume || num_or(volume->get("thicknessFactor"), 0.0) == 0.0;
  double sheen[3] = {0, 0, 0};
  read_vec(jget(jget(ext, "KHR_materials_sheen"), "sheenColorFactor"), sheen, 3);
  double albedo[3] = {bc[0], bc[1], bc[2]};
  double emission[3] = {ef[0] * strength, ef[1] * strength, ef[2] * strength};
  double metallic = num_or(jget(pbr, "metallicFactor"), 1.0);
  int32_t emissive = emissive_map, base = base_map;
  if (jget(ext, "KHR_materials_unlit")) {
    // Unlit: the base colour becomes the emission, nothing reflects.
    for (int k = 0; k < 3; ++k) {
      emission[k] = albedo[k];
      albedo[k] = 0.0;
    }
    emissive = base_map;
    metallic = 0.0;
  }
  return {
      albedo[0], albedo[1], albedo[2], metallic,
      num_or(jget(pbr, "roughnessFactor"), 1.0), num_or(jget(jget(ext, "KHR_materials_ior"), "ior"), 1.5),
      num_or(jget(jget(ext, "KHR_materials_transmission"), "transmissionFactor"), 0.0), 0.0,
      emission[0], emission[1], emission[2], bc[3],
      0.2, 0.2, 0.2, alpha_code,
      double(base), double(mr_map), double(normal_map), double(emissive),
      num_or(jget(infos[2], "scale"), 1.0), num_or(jget(jget(ext, "KHR_materials_clearcoat"), "clearcoatFactor"), 0.0),
      num_or(jget(jget(ext, "KHR_materials_clearcoat"), "clearcoatRoughnessFactor"), 0.0), num_or(jget(m, "alphaCutoff"), 0.5),
      sheen[0], sheen[1], sheen[2], num_or(jget(jget(ext, "KHR_materials_sheen"), "sheenRoughnessFactor"), 0.0),
      att_color[0], att_color[1], att_color[2], std::isfinite(att_dist) ? att_dist : 0.0,
      uv[0], uv[1], uv[2], thin ? 1.0 : 0.0,
      uv[3], uv[4], uv[5], 0.0};
}

// Primitives without a material get the one glTF 2.0 specifies for that
// case (albedo 1, metallic 1, roughness 1, everything else at its default).
std::array<double, 40> default_material() {
  return {1, 1, 1, 1, 1, 1.5, 0, 0, 0, 0, 0, 1, 0.2, 0.2, 0.2, 0, -1, -1, -1, -1,
          1, 0, 0, 0.5, 0, 0, 0, 0, 1, 1, 1, 0, 1, 0, 0, 0, 0, 1, 0, 0};
}

struct Builder {
  Doc* d;
  std::map<long, Accessor> cache;
  long default_slot = -1;

  int32_t material_for(long index) {
    if (index >= 0 && size_t(index) < d->materials.size()) return int32_t(index);
    if (default_slot < 0) {
      default_slot = long(d->materials.size());
      d->materials.push_back(default_material());
    }
    return int32_t(default_slot);
  }

  Status primitive(const Json* prim, const Mat4& world) {
    const long mode = long(num_or(prim->get("mode"), 4.0));
    if (mode != 4 && mode != 5 && mode != 6) {
      warn(d, "point and line primitives are skipped");
      return LUCENT_OK;
    }
    if (jget(jget(prim, "extensions"), "KHR_draco_mesh_compression"))
      return fail(LUCENT_E_UNSUPPORTED, "Draco-compressed meshes are not supported");
    const Json* attrs = prim->get("attributes");
    const long pos_i = idx_or(jget(attrs, "POSITION"));
    if (pos_i < 0) return LUCENT_OK;
    const Accessor *pos = nullptr, *nrm = nullptr, *uv = nullptr, *ind = nullptr;
    Status s = read_accessor(d, pos_i, &cache, &pos);
    if (s != LUCENT_OK) return s;
    if (idx_or(jget(attrs, "NORMAL")) >= 0 && (s = read_accessor(d, idx_or(jget(attrs, "NORMAL")), &cache, &nrm)) != LUCENT_OK) return s;
    if (idx_or(jget(attrs, "TEXCOORD_0")) >= 0 &&
        (s = read_accessor(d, idx_or(jget(attrs, "TEXCOORD_0")), &cache, &uv)) != LUCENT_OK)
      return s;
    if (idx_or(prim->get("indices")) >= 0 && (s = read_accessor(d, idx_or(prim->get("indices")), &cache, &ind)) != LUCENT_OK) return s;
    const size_t nverts = pos->count;
    const size_t n = ind ? ind->data.size() : nverts;
    auto vtx = [&](size_t k) -> double { return ind ? double(ind->data[k]) : double(k); };
    std::vector<double> corners;  // vertex index per corner
    if (mode == 4) {
      for (size_t k = 0; k + 2 < n; k += 3) corners.insert(corners.end(), {vtx(k), vtx(k + 1), vtx(k + 2)});
    } else if (mode == 5) {
      for (size_t k = 0; k + 2 < n; ++k) {
        if (k % 2 == 0) corners.insert(corners.end(), {vtx(k), vtx(k + 1), vtx(k + 2)});
        else corners.insert(corners.end(), {vtx(k), vtx(k + 2), vtx(k + 1)});
      }
    } else {
      for (size_t k = 1; k + 1 < n; ++k) corners.insert(corners.end(), {vtx(k), vtx(k + 1), vtx(0)});
    }
    const size_t tris = corners.size() / 3;
    if (tris == 0) return LUCENT_OK;
    // Normal matrix: the inverse transpose of the upper 3x3 (row-major).
    const double a = world[0], b = world[4], c = world[8], dd = world[1], e = world[5], f = world[9], g = world[2],
                 h = world[6], i = world[10];
    const double A = e * i - f * h, B = -(dd * i - f * g), C = dd * h - e * g;
    const double det = a * A + b * B + c * C;
    const double inv = det == 0 ? 0 : 1 / det;
    const double nm[9] = {A * inv, B * inv, C * inv, -(b * i - c * h) * inv, (a * i - c * g) * inv,
                          -(a * h - b * g) * inv, (b * f - c * e) * inv, -(a * f - c * dd) * inv, (a * e - b * dd) * inv};
    const bool mirror = det < 0;
    Mesh out;
    out.pos.resize(tris * 9);
    out.nrm.resize(tris * 9);
    if (uv) out.uv.resize(tris * 6);
    for (size_t t = 0; t < tris; ++t) {
      for (int k = 0; k < 3; ++k) {
        // det < 0 turns counter-clockwise corners clockwise in world
        // space; reading them 0, 2, 1 restores the front face that the
        // transformed normals point out of.
        const int corner = mirror && k > 0 ? 3 - k : k;
        const double vd = corners[t * 3 + corner];
        if (vd < 0 || vd >= double(nverts))
          return fail(LUCENT_E_INVALID, "index " + std::to_string(long(vd)) + " out of range (" + std::to_string(nverts) + " vertices)");
        const size_t v = size_t(vd);
        const double x = pos->data[v * pos->size], y = pos->data[v * pos->size + 1], z = pos->data[v * pos->size + 2];
        const double wx = world[0] * x + world[4] * y + world[8] * z + world[12];
        const double wy = world[1] * x + world[5] * y + world[9] * z + world[13];
        const double wz = world[2] * x + world[6] * y + world[10] * z + world[14];
        float* po = &out.pos[t * 9 + k * 3];
        po[0] = float(wx);
        po[1] = float(wy);
        po[2] = float(wz);
        const double w3[3] = {wx, wy, wz};
        for (int q = 0; q < 3; ++q) {
          d->bounds[q] = std::min(d->bounds[q], w3[q]);
          d->bounds[q + 3] = std::max(d->bounds[q + 3], w3[q]);
        }
        if (nrm) {
          const double na = nrm->data[v * nrm->size], nb = nrm->data[v * nrm->size + 1], nc = nrm->data[v * nrm->size + 2];
          const double nx = nm[0] * na + nm[1] * nb + nm[2] * nc;
          const double ny = nm[3] * na + nm[4] * nb + nm[5] * nc;
          const double nz = nm[6] * na + nm[7] * nb + nm[8] * nc;
          double len = std::sqrt(nx * nx + ny * ny + nz * nz);
          if (len == 0) len = 1;
          float* no = &out.nrm[t * 9 + k * 3];
          no[0] = float(nx / len);
          no[1] = float(ny / len);
          no[2] = float(nz / len);
        }
        if (uv) {
          out.uv[t * 6 + k * 2] = uv->data[v * uv->size];
          out.uv[t * 6 + k * 2 + 1] = uv->data[v * uv->size + 1];
        }
      }
      if (!nrm) {
        // No normals: flat shading from the stored corners.
        const float* p = &out.pos[t * 9];
        const double e1[3] = {double(p[3]) - p[0], double(p[4]) - p[1], double(p[5]) - p[2]};
        const double e2[3] = {double(p[6]) - p[0], double(p[7]) - p[1], double(p[8]) - p[2]};
        double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        double len = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (len == 0) len = 1;
        for (int k = 0; k < 3; ++k) {
          out.nrm[t * 9 + k * 3] = float(cx / len);
          out.nrm[t * 9 + k * 3 + 1] = float(cy / len);
          out.nrm[t * 9 + k * 3 + 2] = float(cz / len);
        }
      }
    }
    d->triangles += tris;
    out.material = material_for(idx_or(prim->get("material")));
    d->meshes.push_back(std::move(out));
    return LUCENT_OK;
  }

  Status node(long index, const Mat4& parent, int depth) {
    if (depth > 64) return fail(LUCENT_E_INVALID, "node hierarchy is deeper than 64 levels (cycle?)");
    const Json* nodes = d->json.get("nodes");
    const Json* n = nodes ? nodes->at(size_t(index)) : nullptr;
    if (!n || index < 0) return fail(LUCENT_E_INVALID, "missing node " + std::to_string(index));
    Mat4 local;
    if (const Json* mj = n->get("matrix")) {
      for (int k = 0; k < 16; ++k) local[k] = num_or(mj->at(size_t(k)), 0.0);
    } else {
      double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
      read_vec(n->get("translation"), t, 3);
      read_vec(n->get("rotation"), q, 4);
      read_vec(n->get("scale"), s, 3);
      local = trs(t, q, s);
    }
    const Mat4 world = mul4(parent, local);
    const long mesh_i = idx_or(n->get("mesh"));
    if (mesh_i >= 0) {
      const Json* meshes = d->json.get("meshes");
      const Json* mesh = meshes ? meshes->at(size_t(mesh_i)) : nullptr;
      if (!mesh)
        return fail(LUCENT_E_INVALID, "node " + std::to_string(index) + " references missing mesh " + std::to_string(mesh_i));
      if (const Json* prims = mesh->get("primitives"))
        for (const Json& p : prims->arr) {
          Status s = primitive(&p, world);
          if (s != LUCENT_OK) return s;
        }
    }
    if (const Json* ch = n->get("children"))
      for (const Json& c : ch->arr) {
        Status s = node(long(c.num), world, depth + 1);
        if (s != LUCENT_OK) return s;
      }
    return LUCENT_OK;
  }
};

}  // namespace
}  // namespace lucent

using namespace lucent;

extern "C" {

lucent_status lucent_gltf_parse(const uint8_t* bytes, size_t length, lucent_gltf** out) {
  clear_error();
  if (!out || (!bytes && length)) return fail(LUCENT_E_INVALID, "null argument");
  *out = nullptr;
  auto d = std::make_unique<lucent_gltf>();
  const char* json_text = reinterpret_cast<const char*>(bytes);
  size_t json_len = length;
  if (length >= 12 && le32(bytes) == kGlbMagic) {
    // GLB: header, then (length, type, data padded to 4) chunks.
    const uint32_t version = le32(bytes + 4);
    if (version != 2) return fail(LUCENT_E_UNSUPPORTED, "GLB version " + std::to_string(version) + " is not supported; only glTF 2.0");
    const size_t total = std::min<size_t>(le32(bytes + 8), length);
    size_t off = 12;
    json_text = nullptr;
    while (off + 8 <= total) {
      const uint32_t n = le32(bytes + off), type = le32(bytes + off + 4);
      const size_t start = off + 8;
      if (start + n > total) return fail(LUCENT_E_INVALID, "GLB chunk runs past the end of the file");
      if (type == kChunkJson) {
        json_text = reinterpret_cast<const char*>(bytes + start);
        json_len = n;
      } else if (type == kChunkBin && !d->has_glb_bin) {
        d->glb_bin.assign(bytes + start, bytes + start + n);
        d->has_glb_bin = true;
      }
      off = start + n + (4 - n % 4) % 4;
    }
    if (!json_text) return fail(LUCENT_E_INVALID, "GLB has no JSON chunk");
  }
  Status s = JsonReader(json_text, json_len).read(&d->json);
  if (s != LUCENT_OK) return s;
  const Json* ver = jget(jget(&d->json, "asset"), "version");
  std::string version = ver && ver->kind == Json::kStr ? ver->str : ver && ver->is_num() ? std::to_string(ver->num) : "";
  if (version.empty() || version[0] != '2')
    return fail(LUCENT_E_UNSUPPORTED, "glTF version \"" + (version.empty() ? std::string("unknown") : version) +
                                          "\" is not supported; only 2.0");
  if (const Json* req = d->json.get("extensionsRequired"))
    for (const Json& e : req->arr)
      for (const char* r : kRefusedRequired)
        if (e.kind == Json::kStr && e.str == r)
          return fail(LUCENT_E_UNSUPPORTED, "this file requires " + e.str + ", which Lucent does not support");
  // Buffers: the GLB chunk, data URIs, or external resources.
  if (const Json* bufs = d->json.get("buffers")) {
    for (size_t i = 0; i < bufs->arr.size(); ++i) {
      const Json* uri = bufs->arr[i].get("uri");
      d->buffer_bytes.emplace_back();
      d->buffer_resource.push_back(-1);
      if (!uri || uri->kind != Json::kStr) {
        if (i == 0 && d->has_glb_bin) {
          d->buffer_bytes.back() = d->glb_bin;
          continue;
        }
        return fail(LUCENT_E_INVALID, "buffer " + std::to_string(i) + " has no uri and there is no GLB binary chunk");
      }
      if (starts_with(uri->str, "data:")) {
        if ((s = decode_data_uri(uri->str, &d->buffer_bytes.back())) != LUCENT_OK) return s;
      } else {
        d->buffer_resource.back() = add_resource(d.get(), uri->str);
      }
    }
  }
  // Materials (which register the images they sample, in first-use order).
  if (const Json* mats = d->json.get("materials"))
    for (const Json& m : mats->arr) d->materials.push_back(convert_material(d.get(), &m));
  *out = d.release();
  return LUCENT_OK;
}

uint32_t lucent_gltf_resource_count(const lucent_gltf* doc) { return doc ? uint32_t(doc->resources.size()) : 0u; }

const char* lucent_gltf_resource_uri(const lucent_gltf* doc, uint32_t i) {
  return doc && i < doc->resources.size() ? doc->resources[i].uri.c_str() : nullptr;
}

lucent_status lucent_gltf_set_resource(lucent_gltf* doc, uint32_t i, const uint8_t* bytes, size_t length) {
  clear_error();
  if (!doc || (!bytes && length)) return fail(LUCENT_E_INVALID, "null argument");
  if (i >= doc->resources.size()) return fail(LUCENT_E_INVALID, "resource index out of range");
  doc->resources[i].bytes.assign(bytes, bytes + length);
  doc->resources[i].ready = true;
  return LUCENT_OK;
}

lucent_status lucent_gltf_build(lucent_gltf* doc) {
  clear_error();
  if (!doc) return fail(LUCENT_E_INVALID, "null argument");
  if (doc->built) return LUCENT_OK;
  for (const Resource& r : doc->resources)
    if (!r.ready) return fail(LUCENT_E_INVALID, "resource " + r.uri + " was not supplied");
  // Image bytes: a bufferView, a data URI or a supplied resource.
  const Json* imgs = doc->json.get("images");
  for (Image& im : doc->images) {
    const Json* img = imgs ? imgs->at(size_t(im.source)) : nullptr;
    if (im.resource >= 0) {
      im.bytes = doc->resources[size_t(im.resource)].bytes;
    } else if (const Json* bv = jget(img, "bufferView")) {
      const uint8_t* p;
      size_t len, stride;
      Status s = view_bytes(doc, idx_or(bv), &p, &len, &stride);
      if (s != LUCENT_OK) return s;
      im.bytes.assign(p, p + len);
    } else if (const Json* uri = jget(img, "uri")) {
      Status s = decode_data_uri(uri->str, &im.bytes);
      if (s != LUCENT_OK) return s;
    }
  }
  Builder b{doc, {}, -1};
  std::vector<long> roots;
  const long scene = long(num_or(doc->json.get("scene"), 0.0));
  const Json* scenes = doc->json.get("scenes");
  if (scenes && scenes->at(size_t(scene))) {
    if (const Json* ns = scenes->arr[size_t(scene)].get("nodes"))
      for (const Json& n : ns->arr) roots.push_back(long(n.num));
  } else if (const Json* nodes = doc->json.get("nodes")) {
    std::vector<bool> child(nodes->arr.size(), false);
    for (const Json& n : nodes->arr)
      if (const Json* ch = n.get("children"))
        for (const Json& c : ch->arr)
          if (c.num >= 0 && size_t(c.num) < child.size()) child[size_t(c.num)] = true;
    for (size_t i = 0; i < child.size(); ++i)
      if (!child[i]) roots.push_back(long(i));
  }
  for (long r : roots) {
    Status s = b.node(r, identity4(), 0);
    if (s != LUCENT_OK) return s;
  }
  if (doc->triangles == 0) return fail(LUCENT_E_INVALID, "the file contains no triangle meshes");
  doc->built = true;
  return LUCENT_OK;
}

lucent_status lucent_gltf_info(const lucent_gltf* doc, lucent_gltf_summary* out) {
  clear_error();
  if (!doc || !out) return fail(LUCENT_E_INVALID, "null argument");
  if (!doc->built) return fail(LUCENT_E_INVALID, "lucent_gltf_build has not run");
  out->mesh_count = uint32_t(doc->meshes.size());
  out->material_count = uint32_t(doc->materials.size());
  out->image_count = uint32_t(doc->images.size());
  out->warning_count = uint32_t(doc->warnings.size());
  out->triangle_count = doc->triangles;
  for (int k = 0; k < 6; ++k) out->bounds[k] = doc->bounds[k];
  return LUCENT_OK;
}

lucent_status lucent_gltf_mesh(const lucent_gltf* doc, uint32_t i, uint32_t* tri_count, int32_t* material,
                               const float** positions, const float** normals, const float** uvs) {
  clear_error();
  if (!doc || !tri_count || !material || !positions || !normals || !uvs) return fail(LUCENT_E_INVALID, "null argument");
  if (!doc->built || i >= doc->meshes.size()) return fail(LUCENT_E_INVALID, "mesh index out of range");
  const Mesh& m = doc->meshes[i];
  *tri_count = uint32_t(m.pos.size() / 9);
  *material = m.material;
  *positions = m.pos.data();
  *normals = m.nrm.data();
  *uvs = m.uv.empty() ? nullptr : m.uv.data();
  return LUCENT_OK;
}

lucent_status lucent_gltf_material(const lucent_gltf* doc, uint32_t i, double* packed40) {
  clear_error();
  if (!doc || !packed40) return fail(LUCENT_E_INVALID, "null argument");
  if (!doc->built || i >= doc->materials.size()) return fail(LUCENT_E_INVALID, "material index out of range");
  std::memcpy(packed40, doc->materials[i].data(), 40 * sizeof(double));
  return LUCENT_OK;
}

lucent_status lucent_gltf_image(const lucent_gltf* doc, uint32_t i, const uint8_t** data, size_t* length,
                                const char** mime_type, const char** name) {
  clear_error();
  if (!doc || !data || !length || !mime_type || !name) return fail(LUCENT_E_INVALID, "null argument");
  if (!doc->built || i >= doc->images.size()) return fail(LUCENT_E_INVALID, "image index out of range");
  const Image& im = doc->images[i];
  *data = im.bytes.data();
  *length = im.bytes.size();
  *mime_type = im.mime.c_str();
  *name = im.name.c_str();
  return LUCENT_OK;
}

const char* lucent_gltf_warning(const lucent_gltf* doc, uint32_t i) {
  return doc && i < doc->warnings.size() ? doc->warnings[i].c_str() : nullptr;
}

void lucent_gltf_destroy(lucent_gltf* doc) { delete doc; }

}  // extern "C"
