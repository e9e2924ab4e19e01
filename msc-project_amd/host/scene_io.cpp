// scene_io.cpp -- YAML subset, OBJ (tinyobj semantics) and EXR writer.
#include "scene_io.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <stdexcept>
#include <tuple>

namespace swr {
namespace io {

const Node* Node::find(const std::string& key) const {
  if (kind != kMap) return nullptr;
  for (const auto& kv : map)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

namespace {

[[noreturn]] void yaml_error(int line, const std::string& msg) {
  throw std::runtime_error("yaml line " + std::to_string(line) + ": " + msg);
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

// Removes a comment that starts at '#' at line start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '\'' || c == '"') quote = c;
    if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
  }
  return s;
}

Node scalar_node(const std::string& raw) {
  Node n;
  std::string t = trim(raw);
  if (t.empty() || t == "~" || t == "null") {
    n.kind = Node::kNull;
    return n;
  }
  n.kind = Node::kScalar;
  if (t.size() >= 2 && (t[0] == '"' || t[0] == '\'') && t.back() == t[0]) {
    std::string out;
    for (size_t i = 1; i + 1 < t.size(); ++i) {
      if (t[0] == '"' && t[i] == '\\' && i + 2 < t.size()) {
        ++i;
        out += t[i] == 'n' ? '\n' : (t[i] == 't' ? '\t' : t[i]);
      } else if (t[0] == '\'' && t[i] == '\'' && i + 2 < t.size() && t[i + 1] == '\'') {
        out += '\'';
        ++i;
      } else {
        out += t[i];
      }
    }
    n.scalar = out;
  } else {
    n.scalar = t;
  }
  return n;
}

// Flow collections: {k: v, ...} and [a, b, ...], nesting allowed.
struct FlowParser {
  const std::string& s;
  size_t pos = 0;
  int line;

  void ws() {
    while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r')) ++pos;
  }

  std::string plain(const char* stops) {
    ws();
    if (pos < s.size() && (s[pos] == '"' || s[pos] == '\'')) {
      const char q = s[pos];
      size_t end = pos + 1;
      while (end < s.size() && s[end] != q) end += (s[end] == '\\' && q == '"') ? 2 : 1;
      if (end >= s.size()) yaml_error(line, "unterminated quoted scalar");
      std::string out = s.substr(pos, end - pos + 1);
      pos = end + 1;
      return out;
    }
    const size_t start = pos;
    while (pos < s.size() && !std::strchr(stops, s[pos])) ++pos;
    return s.substr(start, pos - start);
  }

  Node value() {
    ws();
    if (pos >= s.size()) yaml_error(line, "unexpected end of flow collection");
    if (s[pos] == '{') {
      ++pos;
      Node n;
      n.kind = Node::kMap;
      ws();
      if (pos < s.size() && s[pos] == '}') {
        ++pos;
        return n;
      }
      for (;;) {
        std::string key = trim(scalar_node(plain(":,}")).scalar);
        ws();
        if (pos >= s.size() || s[pos] != ':') yaml_error(line, "expected ':' after key '" + key + "'");
        ++pos;
        Node v = value();
        n.map.emplace_back(key, std::move(v));
        ws();
        if (pos < s.size() && s[pos] == ',') {
          ++pos;
          continue;
        }
        if (pos < s.size() && s[pos] == '}') {
          ++pos;
          return n;
        }
        yaml_error(line, "expected ',' or '}' in flow mapping");
      }
    }
    if (s[pos] == '[') {
      ++pos;
      Node n;
      n.kind = Node::kSeq;
      ws();
      if (pos < s.size() && s[pos] == ']') {
        ++pos;
        return n;
      }
      for (;;) {
        n.seq.push_back(value());
        ws();
        if (pos < s.size() && s[pos] == ',') {
          ++pos;
          continue;
        }
        if (pos < s.size() && s[pos] == ']') {
          ++pos;
          return n;
        }
        yaml_error(line, "expected ',' or ']' in flow sequence");
      }
    }
    return scalar_node(plain(",]}"));
  }
};

bool balanced(const std::string& s) {
  int depth = 0;
  char quote = 0;
  for (char c : s) {
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    if (c == '{' || c == '[') ++depth;
    if (c == '}' || c == ']') --depth;
  }
  return depth <= 0;
}

struct Line {
  int indent;
  std::string text;  // without indentation
  int number;
};

// Position of the ':' that ends a block mapping key, or npos.
size_t key_colon(const std::string& t) {
  char quote = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    const char c = t[i];
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    if (c == '{' || c == '[') return std::string::npos;
    if (c == ':' && (i + 1 == t.size() || t[i + 1] == ' ' || t[i + 1] == '\t')) return i;
  }
  return std::string::npos;
}

struct BlockParser {
  std::vector<Line> lines;
  size_t i = 0;

  // Value text starting on line idx, extended over following lines while flow brackets are open.
  Node inline_value(std::string text, int number) {
    text = trim(text);
    if (!text.empty() && (text[0] == '{' || text[0] == '[')) {
      while (!balanced(text) && i < lines.size()) {
        text += " " + lines[i].text;
        ++i;
      }
      FlowParser fp{text, 0, number};
      Node n = fp.value();
      fp.ws();
      if (fp.pos != text.size()) yaml_error(number, "trailing text after flow collection");
      return n;
    }
    return scalar_node(text);
  }

  Node block(int indent) {
    Node n;
    if (i >= lines.size()) return n;
    if (lines[i].text[0] == '{' || lines[i].text[0] == '[') {  // a flow collection on its own
      const Line ln = lines[i++];
      return inline_value(ln.text, ln.number);
    }
    const bool seq = lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-";
    n.kind = seq ? Node::kSeq : Node::kMap;
    while (i < lines.size() && lines[i].indent == indent) {
      const Line ln = lines[i];
      if (seq) {
        if (!(ln.text.rfind("- ", 0) == 0 || ln.text == "-")) yaml_error(ln.number, "expected '- ' in sequence");
        ++i;
        std::string rest = ln.text.size() > 1 ? ln.text.substr(2) : "";
        if (trim(rest).empty()) {
          n.seq.push_back(i < lines.size() && lines[i].indent > indent ? block(lines[i].indent) : Node());
        } else if (key_colon(trim(rest)) != std::string::npos) {
          // "- key: value" starts a mapping whose further keys sit at the dash content indent
          const int inner = indent + 2 + (int)(rest.size() - trim(rest).size());
          lines[--i] = Line{inner, trim(rest), ln.number};
          n.seq.push_back(block(inner));
        } else {
          n.seq.push_back(inline_value(rest, ln.number));
        }
        continue;
      }
      const size_t c = key_colon(ln.text);
      if (c == std::string::npos) yaml_error(ln.number, "expected 'key: value'");
      std::string key = scalar_node(ln.text.substr(0, c)).scalar;
      std::string rest = ln.text.substr(c + 1);
      ++i;
      Node v;
      if (trim(rest).empty()) {
        if (i < lines.size() && lines[i].indent > indent) v = block(lines[i].indent);
      } else {
        v = inline_value(rest, ln.number);
      }
      n.map.emplace_back(std::move(key), std::move(v));
    }
    if (i < lines.size() && lines[i].indent > indent) yaml_error(lines[i].number, "bad indentation");
    return n;
  }
};

}  // namespace

std::vector<Node> parse_yaml(const std::string& text) {
  std::vector<Node> docs;
  std::vector<Line> cur;
  bool started = false;
  auto finish = [&]() {
    BlockParser bp{cur, 0};
    Node d;
    if (!cur.empty()) {
      if (cur[0].indent != 0) yaml_error(cur[0].number, "document must start at column 0");
      d = bp.block(0);
      if (bp.i != cur.size()) yaml_error(cur[bp.i].number, "unexpected content");
    }
    docs.push_back(std::move(d));
    cur.clear();
  };
  std::istringstream in(text);
  std::string raw;
  int number = 0;
  while (std::getline(in, raw)) {
    ++number;
    std::string s = strip_comment(raw);
    while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
    if (s.rfind("%", 0) == 0) continue;  // directives
    if (s == "---" || s.rfind("--- ", 0) == 0) {
      if (started || !cur.empty()) finish();
      started = true;
      const std::string rest = trim(s.substr(3));
      if (!rest.empty()) cur.push_back(Line{0, rest, number});
      continue;
    }
    if (s == "...") {
      finish();
      started = false;
      continue;
    }
    if (trim(s).empty()) continue;
    size_t ind = 0;
    while (ind < s.size() && s[ind] == ' ') ++ind;
    if (ind < s.size() && s[ind] == '\t') yaml_error(number, "tabs are not allowed for indentation");
    cur.push_back(Line{(int)ind, s.substr(ind), number});
    started = true;
  }
  if (started || !cur.empty()) finish();
  return docs;
}

// ---- OBJ ------------------------------------------------------------------------------------
namespace {

int fix_index(int idx, int n) {  // tiny_obj_loader.cpp:82-86
  if (idx > 0) return idx - 1;
  if (idx == 0) return 0;
  return n + idx;
}

struct Corner {
  int v = -1, vn = -1, vt = -1;
  bool operator<(const Corner& o) const { return std::tie(v, vn, vt) < std::tie(o.v, o.vn, o.vt); }
};

// "i", "i/j", "i//k", "i/j/k" (tiny_obj_loader.cpp:276-307).
Corner parse_corner(const char*& p, int nv, int nvn, int nvt) {
  Corner c;
  c.v = fix_index(std::atoi(p), nv);
  p += std::strcspn(p, "/ \t\r");
  if (*p != '/') return c;
  ++p;
  if (*p == '/') {
    ++p;
    c.vn = fix_index(std::atoi(p), nvn);
    p += std::strcspn(p, "/ \t\r");
    return c;
  }
  c.vt = fix_index(std::atoi(p), nvt);
  p += std::strcspn(p, "/ \t\r");
  if (*p != '/') return c;
  ++p;
  c.vn = fix_index(std::atoi(p), nvn);
  p += std::strcspn(p, "/ \t\r");
  return c;
}

float parse_float(const char*& p) {
  p += std::strspn(p, " \t");
  const char* end = p + std::strcspn(p, " \t\r");
  std::string tok(p, end);
  p = end;
  return (float)std::strtod(tok.c_str(), nullptr);
}

}  // namespace

ObjMesh load_obj(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open OBJ file " + path);
  std::vector<float> v, vn, vt;
  std::vector<std::vector<Corner>> faces;
  std::string line;
  while (std::getline(f, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
    const char* p = line.c_str();
    p += std::strspn(p, " \t");
    if (*p == '\0' || *p == '#') continue;
    auto space = [](char c) { return c == ' ' || c == '\t'; };
    if (p[0] == 'v' && space(p[1])) {
      p += 2;
      for (int k = 0; k < 3; ++k) v.push_back(parse_float(p));
    } else if (p[0] == 'v' && p[1] == 'n' && space(p[2])) {
      p += 3;
      for (int k = 0; k < 3; ++k) vn.push_back(parse_float(p));
    } else if (p[0] == 'v' && p[1] == 't' && space(p[2])) {
      p += 3;
      for (int k = 0; k < 2; ++k) vt.push_back(parse_float(p));
    } else if (p[0] == 'f' && space(p[1])) {
      p += 2;
      p += std::strspn(p, " \t");
      std::vector<Corner> face;
      while (*p && *p != '\r' && *p != '\n') {
        face.push_back(parse_corner(p, (int)(v.size() / 3), (int)(vn.size() / 3), (int)(vt.size() / 2)));
        p += std::strspn(p, " \t\r");
      }
      faces.push_back(std::move(face));
    }
  }
  // exportFaceGroupToShape (tiny_obj_loader.cpp:346-407): fan triangles, dedup corners.
  ObjMesh m;
  std::map<Corner, uint32_t> cache;
  auto vertex = [&](const Corner& c) -> uint32_t {
    auto it = cache.find(c);
    if (it != cache.end()) return it->second;
    if (c.v < 0 || 3 * (size_t)c.v + 2 >= v.size()) throw std::runtime_error("OBJ vertex index out of range in " + path);
    m.positions.insert(m.positions.end(), {v[3 * c.v], v[3 * c.v + 1], v[3 * c.v + 2]});
    if (c.vn >= 0) {
      if (3 * (size_t)c.vn + 2 >= vn.size()) throw std::runtime_error("OBJ normal index out of range in " + path);
      m.normals.insert(m.normals.end(), {vn[3 * c.vn], vn[3 * c.vn + 1], vn[3 * c.vn + 2]});
    }
    if (c.vt >= 0) {
      if (2 * (size_t)c.vt + 1 >= vt.size()) throw std::runtime_error("OBJ texcoord index out of range in " + path);
      m.texcoords.insert(m.texcoords.end(), {vt[2 * c.vt], vt[2 * c.vt + 1]});
    }
    const uint32_t idx = (uint32_t)(m.positions.size() / 3 - 1);
    cache[c] = idx;
    return idx;
  };
  for (const auto& face : faces) {
    if (face.size() < 3) continue;
    for (size_t k = 2; k < face.size(); ++k) {
      const uint32_t a = vertex(face[0]), b = vertex(face[k - 1]), c = vertex(face[k]);
      m.indices.insert(m.indices.end(), {a, b, c});
    }
  }
  return m;
}

// ---- EXR ------------------------------------------------------------------------------------
uint16_t float_to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xFFu;
  uint32_t man = x & 0x7FFFFFu;
  if (exp == 0xFF) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));  // inf / NaN
  int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7C00u);  // overflow
  if (e <= 0) {                                    // subnormal half or zero
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = man >> shift;
    const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (man >> 13);
  const uint32_t rem = man & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // may carry into the exponent: still correct
  return (uint16_t)(sign | h);
}

float half_to_float(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
  uint32_t x;
  if (exp == 0) {
    if (man == 0) {
      x = sign;
    } else {
      float f = std::ldexp((float)man, -24);
      std::memcpy(&x, &f, 4);
      x |= sign;
    }
  } else if (exp == 31) {
    x = sign | 0x7F800000u | (man << 13);
  } else {
    x = sign | ((exp - 15 + 127) << 23) | (man << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

namespace {

void put_u8(std::vector<uint8_t>& b, uint8_t v) { b.push_back(v); }
void put_i32(std::vector<uint8_t>& b, int32_t v) {
  for (int k = 0; k < 4; ++k) b.push_back((uint8_t)((uint32_t)v >> (8 * k)));
}
void put_u64(std::vector<uint8_t>& b, uint64_t v) {
  for (int k = 0; k < 8; ++k) b.push_back((uint8_t)(v >> (8 * k)));
}
void put_f32(std::vector<uint8_t>& b, float f) {
  int32_t v;
  std::memcpy(&v, &f, 4);
  put_i32(b, v);
}
void put_str(std::vector<uint8_t>& b, const char* s) {
  b.insert(b.end(), s, s + std::strlen(s) + 1);
}
void attr(std::vector<uint8_t>& b, const char* name, const char* type, const std::vector<uint8_t>& value) {
  put_str(b, name);
  put_str(b, type);
  put_i32(b, (int32_t)value.size());
  b.insert(b.end(), value.begin(), value.end());
}

}  // namespace

void write_exr_rgba(const std::string& path, const float* rgb, int width, int height) {
  if (width <= 0 || height <= 0) throw std::runtime_error("exr: empty image");
  std::vector<uint8_t> b;
  put_i32(b, 20000630);  // magic
  put_i32(b, 2);         // version 2, single-part scanline
  // Attributes in name order, as OpenEXR writes them.
  std::vector<uint8_t> ch;
  for (const char* name : {"A", "B", "G", "R"}) {
    put_str(ch, name);
    put_i32(ch, 1);  // HALF
    put_u8(ch, 0);   // pLinear
    put_u8(ch, 0);
    put_u8(ch, 0);
    put_u8(ch, 0);
    put_i32(ch, 1);  // x sampling
    put_i32(ch, 1);  // y sampling
  }
  put_u8(ch, 0);
  attr(b, "channels", "chlist", ch);
  attr(b, "compression", "compression", {0});  // NO_COMPRESSION
  std::vector<uint8_t> box;
  put_i32(box, 0);
  put_i32(box, 0);
  put_i32(box, width - 1);
  put_i32(box, height - 1);
  attr(b, "dataWindow", "box2i", box);
  attr(b, "displayWindow", "box2i", box);
  attr(b, "lineOrder", "lineOrder", {0});  // INCREASING_Y
  std::vector<uint8_t> one;
  put_f32(one, 1.f);
  attr(b, "pixelAspectRatio", "float", one);
  std::vector<uint8_t> centre;
  put_f32(centre, 0.f);
  put_f32(centre, 0.f);
  attr(b, "screenWindowCenter", "v2f", centre);
  attr(b, "screenWindowWidth", "float", one);
  put_u8(b, 0);  // end of header
  const size_t table = b.size();
  const size_t line_bytes = (size_t)width * 4 * 2;
  for (int y = 0; y < height; ++y) put_u64(b, table + 8ull * height + (size_t)y * (8 + line_bytes));
  const uint16_t alpha = float_to_half(1.f);
  for (int y = 0; y < height; ++y) {
    put_i32(b, y);
    put_i32(b, (int32_t)line_bytes);
    for (int c = 0; c < 4; ++c) {  // A, B, G, R
      for (int x = 0; x < width; ++x) {
        const float* p = rgb + 3 * ((size_t)y * width + x);
        const uint16_t h = c == 0 ? alpha : float_to_half(p[3 - c]);
        b.push_back((uint8_t)(h & 0xFF));
        b.push_back((uint8_t)(h >> 8));
      }
    }
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  const int rc = std::fclose(f);
  if (w != b.size() || rc != 0) throw std::runtime_error("short write to " + path);
}

}  // namespace io
}  // namespace swr

namespace swr {
namespace io {
namespace {

constexpr int64_t kMaxTexels = int64_t(1) << 28;

[[noreturn]] void image_error(const std::string& path, const std::string& what) {
  throw std::runtime_error("image " + path + ": " + what);
}

void check_size(const std::string& path, int64_t w, int64_t h) {
  if (w <= 0 || h <= 0 || w * h > kMaxTexels) image_error(path, "size must be positive and at most 2^28 texels");
}

// Header token of a PFM / PPM file: whitespace-separated, '#' comments to the end of the line.
std::string pnm_token(const std::vector<uint8_t>& b, size_t& pos, const std::string& path) {
  for (;;) {
    while (pos < b.size() && std::isspace(b[pos])) ++pos;
    if (pos < b.size() && b[pos] == '#') {
      while (pos < b.size() && b[pos] != '\n') ++pos;
      continue;
    }
    break;
  }
  const size_t start = pos;
  while (pos < b.size() && !std::isspace(b[pos])) ++pos;
  if (pos == start) image_error(path, "truncated header");
  return std::string(b.begin() + start, b.begin() + pos);
}

int64_t pnm_int(const std::vector<uint8_t>& b, size_t& pos, const std::string& path) {
  const std::string t = pnm_token(b, pos, path);
  char* end = nullptr;
  const long long v = std::strtoll(t.c_str(), &end, 10);
  if (*end) image_error(path, "bad header number '" + t + "'");
  return v;
}

Image read_pfm(const std::vector<uint8_t>& b, const std::string& path) {
  size_t pos = 0;
  const std::string magic = pnm_token(b, pos, path);
  if (magic != "PF" && magic != "Pf") image_error(path, "bad PFM magic");
  const int ch = magic == "PF" ? 3 : 1;
  Image im;
  const int64_t w = pnm_int(b, pos, path), h = pnm_int(b, pos, path);
  check_size(path, w, h);
  const std::string st = pnm_token(b, pos, path);
  char* end = nullptr;
  const double scale = std::strtod(st.c_str(), &end);
  if (*end || scale == 0.0) image_error(path, "bad scale '" + st + "'");
  ++pos;  // the single whitespace before the data
  const size_t need = (size_t)(w * h * ch) * 4;
  if (pos > b.size() || b.size() - pos < need) image_error(path, "truncated data");
  const bool little = scale < 0.0;
  im.width = (int)w;
  im.height = (int)h;
  im.rgb.assign((size_t)w * h * 3, 0.f);
  for (int64_t y = 0; y < h; ++y)
    for (int64_t x = 0; x < w; ++x)
      for (int c = 0; c < ch; ++c) {
        const uint8_t* q = &b[pos + 4 * (size_t)((y * w + x) * ch + c)];
        const uint32_t u = little ? (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24
                                  : (uint32_t)q[3] | (uint32_t)q[2] << 8 | (uint32_t)q[1] << 16 | (uint32_t)q[0] << 24;
        float f;
        std::memcpy(&f, &u, 4);
        im.rgb[3 * (size_t)((h - 1 - y) * w + x) + c] = f;  // stored bottom-up
      }
  return im;
}

Image read_ppm(const std::vector<uint8_t>& b, const std::string& path) {
  size_t pos = 0;
  if (pnm_token(b, pos, path) != "P6") image_error(path, "bad PPM magic");
  const int64_t w = pnm_int(b, pos, path), h = pnm_int(b, pos, path), maxval = pnm_int(b, pos, path);
  check_size(path, w, h);
  if (maxval <= 0 || maxval > 65535) image_error(path, "maxval must be in 1..65535");
  ++pos;
  const int bytes = maxval < 256 ? 1 : 2;
  const size_t need = (size_t)(w * h * 3) * bytes;
  if (pos > b.size() || b.size() - pos < need) image_error(path, "truncated data");
  Image im;
  im.width = (int)w;
  im.height = (int)h;
  im.rgb.resize((size_t)w * h * 3);
  for (size_t i = 0; i < im.rgb.size(); ++i) {
    const uint32_t v = bytes == 1 ? b[pos + i] : ((uint32_t)b[pos + 2 * i] << 8 | b[pos + 2 * i + 1]);
    im.rgb[i] = (float)v / (float)maxval;
  }
  return im;
}

template <typename T>
T get_le(const std::vector<uint8_t>& b, size_t pos, const std::string& path) {
  if (pos > b.size() || b.size() - pos < sizeof(T)) image_error(path, "truncated file");
  T v;
  std::memcpy(&v, &b[pos], sizeof(T));
  return v;
}

std::string get_cstr(const std::vector<uint8_t>& b, size_t& pos, const std::string& path) {
  const size_t start = pos;
  while (pos < b.size() && b[pos] != 0) ++pos;
  if (pos >= b.size()) image_error(path, "truncated header");
  return std::string(b.begin() + start, b.begin() + pos++);
}

Image read_exr(const std::vector<uint8_t>& b, const std::string& path) {
  const uint32_t version = get_le<uint32_t>(b, 4, path);
  if ((version & 0xFFu) != 2u || (version & 0x1E00u) != 0u)
    image_error(path, "only single-part scanline EXR (version 2, not tiled, deep or multi-part) is read");
  struct Channel {
    std::string name;
    int32_t type;
  };
  std::vector<Channel> chans;
  int compression = -1;
  int32_t box[4] = {0, 0, -1, -1};
  bool have_box = false;
  size_t pos = 8;
  for (;;) {
    const std::string name = get_cstr(b, pos, path);
    if (name.empty()) break;
    const std::string type = get_cstr(b, pos, path);
    const int32_t size = get_le<int32_t>(b, pos, path);
    pos += 4;
    if (size < 0 || pos + (size_t)size > b.size()) image_error(path, "bad attribute " + name);
    const size_t end = pos + size;
    if (name == "channels") {
      size_t p = pos;
      while (p < end && b[p] != 0) {
        Channel c;
        c.name = get_cstr(b, p, path);
        c.type = get_le<int32_t>(b, p, path);
        const int32_t xs = get_le<int32_t>(b, p + 8, path), ys = get_le<int32_t>(b, p + 12, path);
        if (c.type < 0 || c.type > 2 || xs != 1 || ys != 1) image_error(path, "unsupported channel " + c.name);
        chans.push_back(c);
        p += 16;
      }
    } else if (name == "compression") {
      compression = b[pos];
    } else if (name == "dataWindow") {
      for (int k = 0; k < 4; ++k) box[k] = get_le<int32_t>(b, pos + 4 * k, path);
      have_box = true;
    }
    pos = end;
  }
  if (compression != 0) image_error(path, "only uncompressed EXR is read");
  if (!have_box || chans.empty()) image_error(path, "missing channels or dataWindow");
  const int64_t w = (int64_t)box[2] - box[0] + 1, h = (int64_t)box[3] - box[1] + 1;
  check_size(path, w, h);
  // RGB from channels R, G, B; a lone Y (luminance) is channel 0.
  int slot[3] = {-1, -1, -1};
  for (size_t k = 0; k < chans.size(); ++k) {
    if (chans[k].name == "R") slot[0] = (int)k;
    if (chans[k].name == "G") slot[1] = (int)k;
    if (chans[k].name == "B") slot[2] = (int)k;
  }
  if (slot[0] < 0 && slot[1] < 0 && slot[2] < 0) {
    for (size_t k = 0; k < chans.size(); ++k)
      if (chans[k].name == "Y") slot[0] = (int)k;
    if (slot[0] < 0) image_error(path, "no R, G, B or Y channel");
  }
  std::vector<size_t> ch_off(chans.size());  // byte offset of each channel within a line
  size_t line = 0;
  for (size_t k = 0; k < chans.size(); ++k) {
    ch_off[k] = line;
    line += (size_t)w * (chans[k].type == 1 ? 2 : 4);
  }
  Image im;
  im.width = (int)w;
  im.height = (int)h;
  im.rgb.assign((size_t)w * h * 3, 0.f);
  std::vector<bool> seen((size_t)h, false);
  const size_t table = pos;
  for (int64_t i = 0; i < h; ++i) {
    const uint64_t off = get_le<uint64_t>(b, table + 8 * (size_t)i, path);
    if (off > b.size()) image_error(path, "bad line offset");
    const int64_t y = (int64_t)get_le<int32_t>(b, off, path) - box[1];
    const int32_t size = get_le<int32_t>(b, off + 4, path);
    if (y < 0 || y >= h || seen[(size_t)y] || size < 0 || (size_t)size != line || off + 8 + line > b.size())
      image_error(path, "bad scanline");
    seen[(size_t)y] = true;
    const size_t data = off + 8;
    for (int c = 0; c < 3; ++c) {
      if (slot[c] < 0) continue;
      const Channel& ch = chans[slot[c]];
      for (int64_t x = 0; x < w; ++x) {
        float v;
        if (ch.type == 1) {
          v = half_to_float(get_le<uint16_t>(b, data + ch_off[slot[c]] + 2 * (size_t)x, path));
        } else if (ch.type == 2) {
          v = get_le<float>(b, data + ch_off[slot[c]] + 4 * (size_t)x, path);
        } else {
          v = (float)get_le<uint32_t>(b, data + ch_off[slot[c]] + 4 * (size_t)x, path);
        }
        im.rgb[3 * (size_t)(y * w + x) + c] = v;
      }
    }
  }
  return im;
}

}  // namespace

Image read_image(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open image " + path);
  std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (b.size() >= 4 && get_le<int32_t>(b, 0, path) == 20000630) return read_exr(b, path);
  if (b.size() >= 2 && b[0] == 'P' && (b[1] == 'F' || b[1] == 'f')) return read_pfm(b, path);
  if (b.size() >= 2 && b[0] == 'P' && b[1] == '6') return read_ppm(b, path);
  image_error(path, "unknown format (PFM, binary PPM and uncompressed scanline EXR are read)");
}

}  // namespace io
}  // namespace swr
