// scene_io.cpp -- YAML subset, OBJ (tinyobj semantics), EXR writer and the texture image readers (PFM, PPM, PNG, EXR).
#include "scene_io.h"

#include <algorithm>
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

#include <zlib.h>

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

// PNG (ISO/IEC 15948): chunks checked against their CRC, the IDAT stream inflated with zlib,
// each row unfiltered (None, Sub, Up, Average, Paeth), Adam7 passes placed on the grid. Samples
// become floats as sample / (2^depth - 1) and palette entries as entry / 255 (a division, as the
// PPM reader's sample / maxval). Channels are presented in file
// order (grey: Y; grey + alpha: Y, A; RGB; RGBA; palette: RGB) with missing ones 0, as the other
// readers present them; rows top first.
uint32_t get_be32(const std::vector<uint8_t>& b, size_t pos, const std::string& path) {
  if (pos > b.size() || b.size() - pos < 4) image_error(path, "truncated file");
  return (uint32_t)b[pos] << 24 | (uint32_t)b[pos + 1] << 16 | (uint32_t)b[pos + 2] << 8 | b[pos + 3];
}

Image read_png(const std::vector<uint8_t>& b, const std::string& path) {
  static const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (b.size() < 8 || std::memcmp(b.data(), kSig, 8) != 0) image_error(path, "bad PNG signature");
  int64_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> palette, z;
  bool ended = false;
  for (size_t pos = 8; !ended;) {
    const uint32_t len = get_be32(b, pos, path);
    if (len > 0x7FFFFFFFu || b.size() - pos < 12 + (size_t)len) image_error(path, "truncated chunk");
    const std::string type(b.begin() + pos + 4, b.begin() + pos + 8);
    const uint8_t* data = b.data() + pos + 8;
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), b.data() + pos + 4, 4 + len) != get_be32(b, pos + 8 + len, path))
      image_error(path, "CRC mismatch in chunk " + type);
    if (type == "IHDR") {
      if (len != 13) image_error(path, "bad IHDR");
      w = get_be32(b, pos + 8, path);
      h = get_be32(b, pos + 12, path);
      depth = data[8];
      ctype = data[9];
      interlace = data[12];
      if (data[10] != 0 || data[11] != 0 || interlace > 1) image_error(path, "unknown compression, filter or interlace method");
    } else if (type == "PLTE") {
      palette.assign(data, data + len);
    } else if (type == "IDAT") {
      z.insert(z.end(), data, data + len);
    } else if (type == "IEND") {
      ended = true;
    } else if (!(type[0] & 0x20)) {
      image_error(path, "unknown critical chunk " + type);
    }
    pos += 12 + (size_t)len;
  }
  check_size(path, w, h);
  int chans = 0;
  switch (ctype) {
    case 0: chans = 1; break;
    case 2: chans = 3; break;
    case 3: chans = 1; break;
    case 4: chans = 2; break;
    case 6: chans = 4; break;
    default: image_error(path, "unknown colour type");
  }
  const bool depth_ok = ctype == 0 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)
                        : ctype == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                     : (depth == 8 || depth == 16);
  if (!depth_ok) image_error(path, "bit depth not allowed for the colour type");
  if (ctype == 3 && (palette.empty() || palette.size() % 3 != 0 || palette.size() > 768)) image_error(path, "bad palette");
  // Adam7 passes (x0, y0, dx, dy), or the whole image as one pass.
  static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  static const int kWhole[1][4] = {{0, 0, 1, 1}};
  const int npass = interlace ? 7 : 1;
  const int (*passes)[4] = interlace ? kAdam7 : kWhole;
  const size_t bits = (size_t)chans * depth;
  const size_t bpp = std::max<size_t>(1, bits / 8);
  size_t raw_size = 0;
  for (int p = 0; p < npass; ++p) {
    const int64_t pw = w > passes[p][0] ? (w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0;
    const int64_t ph = h > passes[p][1] ? (h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0;
    if (pw && ph) raw_size += (size_t)ph * (1 + ((size_t)pw * bits + 7) / 8);
  }
  std::vector<uint8_t> raw(raw_size);
  uLongf got = (uLongf)raw_size;
  if (uncompress(raw.data(), &got, z.data(), (uLong)z.size()) != Z_OK || got != raw_size)
    image_error(path, "corrupt or short image data");
  Image im;
  im.width = (int)w;
  im.height = (int)h;
  im.rgb.assign((size_t)w * h * 3, 0.f);
  const float top = (float)((1u << depth) - 1u);
  size_t off = 0;
  for (int p = 0; p < npass; ++p) {
    const int64_t pw = w > passes[p][0] ? (w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0;
    const int64_t ph = h > passes[p][1] ? (h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0;
    if (!pw || !ph) continue;
    const size_t rowb = ((size_t)pw * bits + 7) / 8;
    std::vector<uint8_t> prev(rowb, 0), cur(rowb);
    for (int64_t y = 0; y < ph; ++y) {
      const uint8_t filter = raw[off];
      const uint8_t* in = raw.data() + off + 1;
      off += 1 + rowb;
      for (size_t i = 0; i < rowb; ++i) {
        const int a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
        int pred = 0;
        switch (filter) {
          case 0: pred = 0; break;
          case 1: pred = a; break;
          case 2: pred = up; break;
          case 3: pred = (a + up) / 2; break;
          case 4: {
            const int pa = std::abs(up - c), pb = std::abs(a - c), pc = std::abs(a + up - 2 * c);
            pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
            break;
          }
          default: image_error(path, "unknown row filter");
        }
        cur[i] = (uint8_t)(in[i] + pred);
      }
      const int64_t yy = passes[p][1] + y * passes[p][3];
      for (int64_t x = 0; x < pw; ++x) {
        const int64_t xx = passes[p][0] + x * passes[p][2];
        float* dst = &im.rgb[3 * (size_t)(yy * w + xx)];
        for (int c = 0; c < chans; ++c) {
          const size_t bit = ((size_t)x * chans + c) * depth;
          uint32_t v;
          if (depth == 16)
            v = (uint32_t)cur[bit / 8] << 8 | cur[bit / 8 + 1];
          else if (depth == 8)
            v = cur[bit / 8];
          else
            v = (cur[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
          if (ctype == 3) {
            if (3 * (size_t)v + 2 >= palette.size()) image_error(path, "palette index out of range");
            for (int k = 0; k < 3; ++k) dst[k] = (float)palette[3 * v + k] / 255.f;
          } else if (c < 3) {
            dst[c] = (float)v / top;
          }
        }
      }
      std::swap(prev, cur);
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
  if (b.size() >= 8 && b[0] == 137 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G') return read_png(b, path);
  image_error(path, "unknown format (PFM, binary PPM, PNG and uncompressed scanline EXR are read)");
}

}  // namespace io
}  // namespace swr
